"""Gradient-step rules for design updates.

Public interface of beamz/optimization/optimizers.py:66-107: ``Optimizer(method,
learning_rate, options)``, whose ``step(gradient)`` returns the increment to add
to the design variables (a descent step: the minus sign is inside) and whose
``reset()`` forgets the accumulated state.

The one rule is Adam (Kingma and Ba, 2015): exponential moving averages of the
gradient and of its square, each divided by its bias correction 1 - beta^t,
increment = -lr * mean_hat / (sqrt(square_hat) + epsilon), with the published
defaults beta1 = 0.9, beta2 = 0.999, epsilon = 1e-8 (the ones the reference
uses). ``options`` overrides them by name.
"""
from __future__ import annotations

import numpy as np


class AdamRule:
    """Adam's state for one array of design variables."""

    name = "adam"

    def __init__(self, beta1=0.9, beta2=0.999, epsilon=1e-8):
        self.beta1, self.beta2, self.epsilon = float(beta1), float(beta2), float(epsilon)
        self.reset()

    def reset(self):
        self.mean = None        # moving average of the gradient
        self.square = None      # moving average of its elementwise square
        self.count = 0

    def increment(self, g, lr):
        if self.mean is None or self.mean.shape != g.shape:
            self.mean = np.zeros_like(g)
            self.square = np.zeros_like(g)
            self.count = 0
        self.count += 1
        b1, b2 = self.beta1, self.beta2
        self.mean = b1 * self.mean + (1.0 - b1) * g
        self.square = b2 * self.square + (1.0 - b2) * np.square(g)
        mean_hat = self.mean / (1.0 - b1 ** self.count)
        square_hat = self.square / (1.0 - b2 ** self.count)
        return -lr * mean_hat / (np.sqrt(square_hat) + self.epsilon)


RULES = {rule.name: rule for rule in (AdamRule,)}


class Optimizer:
    """Keeps one update rule's state across iterations; ``step`` maps a gradient to the increment."""

    def __init__(self, method="adam", learning_rate=1e-2, options=None):
        rule = RULES.get(str(method).lower())
        if rule is None:
            raise ValueError(f"optimizer {method!r} is not available; choose one of {sorted(RULES)}")
        self.method = method
        self.learning_rate = learning_rate
        self.options = dict(options or {})
        self.rule = rule(**self.options)

    def step(self, gradient):
        if not isinstance(gradient, np.ndarray):
            raise TypeError(f"the gradient must be a numpy array, not {type(gradient).__name__}")
        if gradient.size == 0:
            return np.zeros_like(gradient)
        return self.rule.increment(gradient, self.learning_rate)

    def reset(self):
        self.rule.reset()


__all__ = ["Optimizer", "AdamRule"]
