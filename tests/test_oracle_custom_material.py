"""Spatially varying materials (the reference's CustomMaterial, beamz/design/materials.py:14-217) in the
rasterizer (beamz/simulation/meshing.py:162-169 and 209-247, axis-aligned rectangle fast path; background and
3D read the material once at (0, 0, 0)).

The oracle's restatement (oracle/raster_ref.py) is pinned first: a constant function rasterizes exactly like
the plain Material of that value; interior cells of a varying function take the function at their cell centre
exactly; a grid material's interior cells equal scipy's RegularGridInterpolator at the centres; a two-argument
function, called with z = 0 as the reference does, fails and gives the defaults (1, 1, 0); edge cells blend
the material at their own centre in ninths. Then the native rasterizer must give the oracle's bits."""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

import beamz_amd as bz
from beamz_amd.design.materials import CustomMaterial
from beamz_amd.simulation.meshing import RegularGrid
from oracle import raster_ref as RR

UM = 1e-6


def design(material, bg=None, pos=(1.1 * UM, 0.93 * UM), size=(3.05 * UM, 2.2 * UM), W=6 * UM, H=4 * UM):
    d = bz.Design(W, H, material=bg or bz.Material(2.0), auto_pml=False)
    d += bz.Rectangle(position=pos, width=size[0], height=size[1], material=material)
    return d


def f3(x, y, z):
    return 2.0 + np.sin(x / UM) * np.cos(0.7 * y / UM) + z


def cond3(x, y, z):
    return 1e3 * (x / UM) + 50.0 * (y / UM)


def centres(W, H, res):
    gw, gh = int(W / res), int(H / res)
    return np.linspace(0.5 * res, W - 0.5 * res, gw), np.linspace(0.5 * res, H - 0.5 * res, gh)


def test_constant_function_equals_the_plain_material():
    res = 0.1 * UM
    a = RR.rasterize_2d(design(CustomMaterial(permittivity_func=lambda x, y, z: 7.5,
                                              conductivity_func=lambda x, y, z: 40.0)), res)
    b = RR.rasterize_2d(design(bz.Material(7.5, 1.0, 40.0)), res)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_interior_cells_take_the_function_at_their_centre_and_edges_blend_it():
    res = 0.1 * UM
    P, M, C = RR.rasterize_2d(design(CustomMaterial(permittivity_func=f3, conductivity_func=cond3)), res)
    xc, yc = centres(6 * UM, 4 * UM, res)
    # interior rows/columns of the fast path: int((p + 0.25 dx)/dx) .. floor((p + w - 0.25 dx)/dx)
    n0j, n1j = int((1.1 * UM + 0.25 * res) / res), int(np.floor((4.15 * UM - 0.25 * res) / res))
    n0i, n1i = int((0.93 * UM + 0.25 * res) / res), int(np.floor((3.13 * UM - 0.25 * res) / res))
    for i in (n0i, (n0i + n1i) // 2, n1i - 1):
        for j in (n0j, (n0j + n1j) // 2, n1j - 1):
            assert P[i, j] == f3(xc[j], yc[i], 0) and C[i, j] == cond3(xc[j], yc[i], 0), (i, j)
    assert np.all(M == 1.0)
    # the top edge row (below the fast path's interior, above the rectangle's top at 3.13 um) blends the
    # value at its centre with the background in ninths
    i, j = n1i, (n0j + n1j) // 2
    y0, y1 = 0.93 * UM, 3.13 * UM
    cnt = sum(3 for oy in (-0.25, 0, 0.25) if y0 <= yc[i] + oy * res < y1)
    assert 0 < cnt < 9
    bf = cnt / 9
    assert P[i, j] == 2.0 * (1 - bf) + f3(xc[j], yc[i], 0) * bf


def test_grid_material_interior_equals_the_interpolator():
    res = 0.1 * UM
    rng = np.random.default_rng(3)
    g = 2.0 + 10.0 * rng.random((17, 23))
    bounds = ((1.5 * UM, 4.3 * UM), (0.8 * UM, 3.3 * UM))
    m = CustomMaterial(permittivity_grid=g, bounds=bounds)
    P, _, _ = RR.rasterize_2d(design(m), res)
    xc, yc = centres(6 * UM, 4 * UM, res)
    interp = RegularGridInterpolator((np.linspace(0.8 * UM, 3.3 * UM, 17), np.linspace(1.5 * UM, 4.3 * UM, 23)), g,
                                     method="linear", bounds_error=False, fill_value=1.0)
    n0j, n1j = int((1.1 * UM + 0.25 * res) / res), int(np.floor((4.15 * UM - 0.25 * res) / res))
    n0i, n1i = int((0.93 * UM + 0.25 * res) / res), int(np.floor((3.13 * UM - 0.25 * res) / res))
    Y, X = np.meshgrid(yc[n0i:n1i], xc[n0j:n1j], indexing="ij")
    want = interp(np.column_stack([Y.ravel(), X.ravel()])).reshape(Y.shape)
    assert np.array_equal(P[n0i:n1i, n0j:n1j], want)
    # outside the grid's bounds the interpolator's fill value, the default permittivity 1
    mid = (n0i + n1i) // 2
    assert xc[n0j] < 1.5 * UM and P[mid, n0j] == 1.0 and P[mid, n1j - 1] != 1.0


def test_two_argument_function_fails_with_z_and_gives_defaults():
    res = 0.1 * UM
    a = RR.rasterize_2d(design(CustomMaterial(permittivity_func=lambda x, y: 9.0)), res)
    b = RR.rasterize_2d(design(bz.Material(1.0, 1.0, 0.0)), res)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_background_reads_the_material_at_the_origin():
    res = 0.1 * UM
    bg = CustomMaterial(permittivity_func=lambda x, y, z: 3.0 + x + y + z)
    P, _, _ = RR.rasterize_2d(design(bz.Material(5.0), bg=bg), res)
    assert P[0, 0] == 3.0 and P[-1, -1] == 3.0


CASES = {
    "function": lambda: CustomMaterial(permittivity_func=f3, conductivity_func=cond3),
    "grid": lambda: CustomMaterial(permittivity_grid=2.0 + 10.0 * np.random.default_rng(3).random((17, 23)),
                                   conductivity_grid=np.random.default_rng(4).random((5, 6)) * 1e3,
                                   bounds=((1.0 * UM, 4.3 * UM), (0.8 * UM, 3.3 * UM))),
    "grid-nearest": lambda: CustomMaterial(permittivity_grid=np.arange(30.0).reshape(5, 6) + 1,
                                           bounds=((1.2 * UM, 3.9 * UM), (1.0 * UM, 2.9 * UM)),
                                           interpolation="nearest"),
    "two-argument": lambda: CustomMaterial(permittivity_func=lambda x, y: 9.0),
}


@pytest.mark.parametrize("kind", sorted(CASES))
@pytest.mark.parametrize("res", [0.1 * UM, 0.0737 * UM])
def test_native_rasterizer_matches_the_oracle(kind, res):
    d = design(CASES[kind](), bg=CustomMaterial(permittivity_func=lambda x, y, z: 2.1))
    d += bz.Rectangle(position=(0.4 * UM, 2.2 * UM), width=1.3 * UM, height=0.37 * UM, material=bz.Material(6.0))
    g = RegularGrid(d, res)
    P, M, C = RR.rasterize_2d(d, res)
    assert np.array_equal(g.permittivity, P)
    assert np.array_equal(g.permeability, M)
    assert np.array_equal(g.conductivity, C)


def test_varying_material_on_other_shapes_is_refused():
    d = bz.Design(6 * UM, 4 * UM, material=bz.Material(2.0), auto_pml=False)
    d += bz.Circle(position=(3 * UM, 2 * UM), radius=1 * UM, material=CustomMaterial(permittivity_func=f3))
    with pytest.raises(NotImplementedError):
        RegularGrid(d, 0.1 * UM)
    with pytest.raises(NotImplementedError):
        RR.rasterize_2d(d, 0.1 * UM)


def test_custom_material_api():
    g = np.arange(12.0).reshape(3, 4)
    m = CustomMaterial(permittivity_grid=g, bounds=((0, 3.0), (0, 2.0)))
    assert m.permittivity == "grid(0.000-11.000)"
    assert float(m.get_permittivity(1.0, 1.0)[0]) == 5.0           # row y = 1 of 0, 1, 2; column x = 1
    m2 = m.copy()
    m2.update_grid("permittivity", g * 2)
    assert float(m2.get_permittivity(1.0, 1.0)[0]) == 10.0 and float(m.get_permittivity(1.0, 1.0)[0]) == 5.0
    f = CustomMaterial(permeability_func=lambda x, y: x + y)
    assert f.permeability == "function" and f.get_permeability(1.0, 2.0) == 3.0
    assert f.get_sample(1.0, 2.0) == (1.0, 3.0, 0.0)
    with pytest.raises(ValueError):
        m.update_grid("refractive_index", g)
