"""PhysicalModel derived scalars: known answers of the reference's own test
(pymcac/tests/test_read.py:27-52 for validation/params_pytest.ini), for the
oracle and for the product's host code."""
import pathlib

import pytest

import oracle_lib as O

PARAMS_PYTEST = """
[environment]
initial_time = 0
fractal_dimension=1.78
fractal_prefactor=1.30
pressure=101300
temperature=1700
volume_fraction=10e-6

[limits]
mean_monomere_per_aggregate=40

[monomers]
density=1800
dispersion_diameter=1.20
initialisation_mode=lognormal
mean_diameter=10
number=20

[output]
n_time_per_file=5000
output_dir=pytest_data

[surface_growth]
with_surface_reactions = true
flux_surfgrowth=1e-04
volsurf_method=alphas
full_aggregate_update_frequency=100
"""
# values as the reference writes them (6 significant digits)
KNOWN = {"box_length": 1.06741e-06, "lambda": 4.98113e-07, "viscosity": 5.662e-05,
         "rpeqmass": 5.25563e-09, "gamma": 1.378, "u_sg": 5.55556e-08}


def sig6(x):
    return float(f"{x:.6g}")


def test_oracle_known_answers():
    got = O.physics(PARAMS_PYTEST)
    for k, v in KNOWN.items():
        assert sig6(got[k]) == v, k


def test_product_host_physics_equals_oracle():
    mcac_amd = pytest.importorskip("mcac_amd")
    try:
        got = mcac_amd.physics_scalars(PARAMS_PYTEST)
    except RuntimeError as e:  # library not built
        pytest.fail(str(e))
    ref = O.physics(PARAMS_PYTEST)
    assert got == ref
    for k, v in KNOWN.items():
        assert sig6(got[k]) == v, k


@pytest.mark.parametrize("ini", sorted((pathlib.Path(__file__).parent / "configs").glob("*.ini")),
                         ids=lambda p: p.stem)
def test_product_physics_equals_oracle_on_configs(ini):
    import mcac_amd
    text = ini.read_text()
    assert mcac_amd.physics_scalars(text) == O.physics(text)


def test_c1_box_length():
    """(N*pi/6*Dpm^3/FV)^(1/3) for the N=1000, 30 nm, FV=1e-3 monodisperse case."""
    text = (pathlib.Path(__file__).parent / "configs" / "c1_cpu_reference.ini").read_text()
    L = O.physics(text)["box_length"]
    import math
    assert L == pytest.approx((1000 * math.pi / 6 * (30e-9) ** 3 / 1e-3) ** (1 / 3), rel=1e-14)
