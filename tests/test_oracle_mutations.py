"""The oracle's pins must be able to fail.  Each case below copies
oracle/rk_oracle.c into a temporary directory, plants one plausible
misreading of the reference in the copy, builds it with the oracle's own
Makefile and routes the oracle to it; the named known-answer checks must
then fail.  The repository's oracle is never modified.

| planted mistake | reference it contradicts | checks that catch it |
|---|---|---|
| Davis speeds at a dry front | rki_kernel_flux.F90:456-467 | dry-bed closed form, Ritter cells |
| THIN_LAYER_THRESHOLD = 1e-6 | rki_constants.F90:59, rki_kernel_flux.F90:158-181 | thin ramp, Ritter cells |
| minmod -> 0 (first order) | rki_kernel_flux.F90:380-393 | thick ramp, observed order 2 |
| HLLC upwinds on -s* | rki_kernel_flux.F90:676-681 | HLLC contact value |
| no hydrostatic pressure correction | rki_kernel_flux.F90:358-369 | lake at rest over a bump |
| Flather sign flipped | rki_boundary.F90:331-343 | Flather sign on every edge |
"""
import shutil
import subprocess
from pathlib import Path

import pytest

import oracle as O

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle"

import test_oracle_kat as KAT  # noqa: E402
import test_oracle_wetdry as WD  # noqa: E402

MUTATIONS = {
    "davis_dry_front": [
        ("s_r = u_l + 2.0 * a_l;", "s_r = u_l + a_l;"),
        ("s_l = u_r - 2.0 * a_r;", "s_l = u_r - a_r;"),
        ("s_r = v_b + 2.0 * a_b;", "s_r = v_b + a_b;"),
        ("s_l = v_t - 2.0 * a_t;", "s_l = v_t - a_t;"),
    ],
    "thin_threshold_1e-6": [
        ("#define THIN_LAYER_THRESHOLD 1.0e-3", "#define THIN_LAYER_THRESHOLD 1.0e-6"),
    ],
    "first_order": [
        ("    if (a * b <= 0.0) return 0.0;\n    if (fabs(a) < fabs(b)) return a;",
         "    if (a * b <= 0.0 || 1) return 0.0;\n    if (fabs(a) < fabs(b)) return a;"),
    ],
    "hllc_upwind_flipped": [
        ("*f_hv = (s_star >= 0.0) ? *f_h * v_l : *f_h * v_r;",
         "*f_hv = (s_star < 0.0) ? *f_h * v_l : *f_h * v_r;"),
    ],
    "flather_sign_flipped": [
        ("+ sgn * speed * (eta_int - eta_ext) / hmax", "- sgn * speed * (eta_int - eta_ext) / hmax"),
    ],
    "no_pressure_correction": [
        ("h_hyd_left = h_hyd_r;", "h_hyd_left = h_R;"),
        ("h_hyd_right = h_hyd_l;", "h_hyd_right = h_L;"),
    ],
}


def build_mutant(tmp: Path, name: str) -> Path:
    d = tmp / name
    d.mkdir()
    for f in ("rk_oracle.c", "rk_oracle.h", "rk_crmath.h"):
        shutil.copy(ORACLE / f, d / f)
    src = (d / "rk_oracle.c").read_text()
    for old, new in MUTATIONS[name]:
        n = src.count(old)
        assert n >= 1, f"{name}: pattern not found: {old!r}"
        src = src.replace(old, new)
    (d / "rk_oracle.c").write_text(src)
    subprocess.run(["make", "-s", "-C", str(d), "-f", str(ORACLE / "Makefile"), "liborc.so"], check=True)
    return d / "liborc.so"


def fails(check) -> bool:
    try:
        check()
    except AssertionError:
        return True
    return False


def conv_series():
    return [WD.standing_mode_level(nx) for nx in (50, 100, 200, 400)]


CATCHERS = {
    "davis_dry_front": {
        "dry-bed closed form (x, HLL)": lambda: WD.test_dry_front_flux_closed_form(0, 0, True),
        "dry-bed closed form (y, HLLC, dry left)": lambda: WD.test_dry_front_flux_closed_form(1, 1, False),
        "Ritter front cells": lambda: WD.test_ritter_front_positions_match_reference(WD.run_ritter()),
    },
    "thin_threshold_1e-6": {
        "thin ramp (x)": lambda: WD.test_thin_layer_ramp_is_first_order(0),
        "thin ramp (y)": lambda: WD.test_thin_layer_ramp_is_first_order(1),
        "Ritter front cells": lambda: WD.test_ritter_front_positions_match_reference(WD.run_ritter()),
    },
    "first_order": {
        "thick ramp": lambda: WD.test_thick_ramp_is_second_order(0),
        "observed order 2": lambda: WD.test_convergence_order_two(conv_series()),
    },
    "hllc_upwind_flipped": {
        "HLLC contact value": lambda: KAT.test_hllc_contact_upwinding_value(),
    },
    "flather_sign_flipped": {
        "Flather sign on every edge": lambda: KAT.test_open_inflow_discharge_clamped(),
    },
    "no_pressure_correction": {
        "lake at rest over a bump (HLL)": lambda: KAT.test_lake_at_rest_zero_flux(0),
    },
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(tmp_path, name):
    lib = build_mutant(tmp_path, name)
    with O.using(lib):
        missed = [label for label, check in CATCHERS[name].items() if not fails(check)]
    assert not missed, f"{name}: these checks pass on the mutant: {missed}"
    # and the same checks pass on the oracle itself
    for label, check in CATCHERS[name].items():
        check()
