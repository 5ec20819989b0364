"""Pins of RK_CHAN_1's explicit right-hand side (tests/oracle_pins.py) on the
oracle, and proof that they bite: the oracle is rebuilt once per planted
mistake (diablo_oracle.c's PLANT sites: a flipped sign, a shifted index, a
swapped weight or operand in one term) and every such copy must be rejected
by the pins listed for it.  The table records which pins catch each
mistake; every term of for/channel.f:123-290 and 410-720 has at least one
entry, as do the Crank-Nicolson halves (606-626, 740-746) and the CP update
(1059-1065)."""
import os
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_pins as op

PLANTED = {
    1: ("CR1 -= H_BAR i kx CP: sign (for/channel.f:136)", ["pressure_projection"]),
    2: ("dCP/dy at GY divides by DY(J), not DY(J+1) (for/channel.f:127)", ["pressure_projection"]),
    3: ("CR3 -= H_BAR i kz CP, not i kx (for/channel.f:137)", ["pressure_projection"]),
    4: ("F_TYPE = 2: AMP_OMEGA0 cos(OMEGA0 TIME), not sin (for/channel.f:163-164)", ["mean_forcing"]),
    5: ("horizontal viscosity (KX2+KZ2)**BETA, not *BETA (for/channel.f:178-179)", ["linear_terms"]),
    6: ("Coriolis CF1 += I_RO CU3: sign (for/channel.f:196)", ["linear_terms"]),
    7: ("Coriolis CF3 -= I_RO CU1: sign (for/channel.f:197)", ["linear_terms"]),
    8: ("buoyancy on GY: CTH(J) weighted by DYF(J-1) (for/channel.f:213-214, 224-225)", ["buoyancy"]),
    9: ("buoyancy skips the (0,0) mode (for/channel.f:217-223)", ["buoyancy"]),
    10: ("RI GRAV_X TH into CF1, not GRAV_Z (for/channel.f:232)", ["linear_terms"]),
    11: ("scalar diffusivity NU/PR (for/channel.f:277)",
         ["linear_terms", "v_products_uniform", "horizontal_products"]),
    12: ("CFTH -= CU1 DRHODX: sign (for/channel.f:286)", ["linear_terms"]),
    13: ("CFTH -= CU3 DRHODZ, not CU1 (for/channel.f:287)", ["linear_terms"]),
    14: ("U1*U3 enters CF1 with i kz (for/channel.f:423)", ["v_products_uniform", "horizontal_products"]),
    15: ("U3*U3 enters CF3 with -i kz: sign (for/channel.f:462)", ["v_products_uniform", "horizontal_products"]),
    16: ("U1 on GY: U1(J) weighted by DYF(J) (for/channel.f:472-473)", ["v_products_stretched"]),
    17: ("U3 on GY: U3(J) weighted by DYF(J) (for/channel.f:493-494)", ["v_products_stretched"]),
    18: ("d/dy(U1 U2): the term U1(J) U2(J+1) (for/channel.f:515)",
         ["v_products_stretched", "v_products_uniform", "free_slip_conservation"]),
    19: ("d/dy(U3 U2) divides by 2 DYF(J) (for/channel.f:537)",
         ["v_products_stretched", "free_slip_conservation"]),
    20: ("d/dy(V^2) at GY divides by DY(J) (for/channel.f:558)", ["v_products_stretched"]),
    21: ("U1*TH enters CFTH with i kx (for/channel.f:681)",
         ["v_products_stretched", "v_products_uniform", "horizontal_products"]),
    22: ("U3*TH enters CFTH with -i kz: sign (for/channel.f:697)",
         ["v_products_stretched", "horizontal_products"]),
    23: ("d/dy(TH U2): the term TH(J-1) U2(J) (for/channel.f:709)",
         ["v_products_stretched", "v_products_uniform", "free_slip_conservation"]),
    24: ("explicit CN half of U1: (U1(J+1)-U1(J))/DY(J+1) (for/channel.f:610)",
         ["free_slip_conservation", "mean_forcing", "cn_mean_velocity", "cn_stokes_mode"]),
    25: ("explicit CN half of U2: (U2(J)-U2(J-1))/DYF(J-1) (for/channel.f:623)", ["cn_stokes_mode"]),
    26: ("explicit CN half of TH: TEMP1/PR (for/channel.f:742)", ["cn_scalar"]),
    27: ("explicit CN half of U1 carries NU_V_SCALE (for/channel.f:609)", ["cn_mean_velocity", "cn_stokes_mode"]),
    28: ("CP += phi / H_BAR, not / (BETA_BAR H_BAR) (for/channel.f:1062)", ["pressure_projection"]),
    29: ("F_TYPE = 1: CR1(0,0) -= H_BAR PX0: sign (for/channel.f:158)", ["mean_forcing"]),
    30: ("U1*U1 enters CF1 with i kx (for/channel.f:443)", ["v_products_uniform", "horizontal_products"]),
}


def _maker(oracle_mod, library=None):
    return lambda cfg, gy_1=None: oracle_mod.Oracle(cfg, gy_1=gy_1, library=library)


@pytest.mark.parametrize("name", list(op.PINS))
def test_pin_holds_on_the_oracle(oracle_mod, name):
    worst = op.PINS[name](_maker(oracle_mod))
    assert worst < 1e-11


@pytest.fixture(scope="module")
def planted(oracle_mod, tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")

    def one(n):
        return n, oracle_mod.build(plant=n, out=str(d / f"libdiablo_oracle_plant{n}.so"))

    with ThreadPoolExecutor(max(1, min(8, os.cpu_count() or 1))) as ex:
        return dict(ex.map(one, sorted(PLANTED)))


@pytest.mark.parametrize("n", sorted(PLANTED))
def test_planted_mistake_is_rejected(oracle_mod, planted, n):
    what, pins = PLANTED[n]
    make = _maker(oracle_mod, planted[n])
    for name in pins:
        with pytest.raises(op.PinFailure):
            op.PINS[name](make)


def test_every_term_has_a_pin():
    assert set(PLANTED) == set(range(1, 31))
    for n, (what, pins) in PLANTED.items():
        assert pins and set(pins) <= set(op.PINS), n
    # every pin is needed by at least one planted mistake
    used = {p for _, pins in PLANTED.values() for p in pins}
    assert used == set(op.PINS)
