"""The flux checks of test_oracle_flux_pins.py must fail on a broken oracle.

Each case plants one plausible transcription mistake in a copy of the oracle's
flux code (oracle/laws.hpp) -- swapped Roe wave strengths, a dropped absolute
value of one eigenvalue, a dropped shear term, the arithmetic instead of the
Roe-weighted mean of the sound speeds, a wrong Matrix eigenvalue scale, a wrong
Kennedy-Gruber gravity weight or mass flux -- builds it, and runs the flux
checks against it in a subprocess; they must fail.  The same build of the
unmodified sources must pass them.  CPU only; the oracle itself is never
changed.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CXX = ["g++", "-O1", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared"]

# (name, text, replacement, occurrence: 1 = Euler, 2 = EulerGravity, 3 = ShallowWater)
MUTANTS = [
    ("roe_swap_w1_w2",
     "double w1 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;\n"
     "    double w2 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;",
     "double w1 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;\n"
     "    double w2 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;", 1),
    ("roe_drop_abs_minus", "std::fabs(un - c) * (dp - rho", "(un - c) * (dp - rho", 1),
    ("roe_gravity_drop_abs_plus", "std::fabs(un + c) * (dp + rho", "(un + c) * (dp + rho", 2),
    ("roe_drop_shear_energy", "w4 * (dotv<D>(u, du) - un * dun)) / 2;", "0.0) / 2;", 1),
    ("roe_arithmetic_sound_speed", "double c = roe_avg(rm, rp, cm, cp);", "double c = avg(cm, cp);", 1),
    ("roe_sw_swap_w1_w2",
     "double w1 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;\n"
     "  double w2 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;",
     "double w1 = std::fabs(un + c) * (dp + rho * c * dun) * cm2 / 2;\n"
     "  double w2 = std::fabs(un - c) * (dp - rho * c * dun) * cm2 / 2;", 1),
    ("roe_sw_tracer_wave", "double w5 = std::fabs(un) * (drt - th * dp * cm2);", "double w5 = std::fabs(un) * drt;", 1),
    ("matrix_drop_abs_l1", "double l1 = std::fabs(uavgn - c_bar)", "double l1 = (uavgn - c_bar)", 1),
    ("matrix_swap_l1_l3", "double l1 = std::fabs(uavgn - c_bar)", "double l1 = std::fabs(uavgn + c_bar)", 1),
    ("matrix_gravity_l2_scale", "double l2 = std::fabs(uavgn) * rho_log * (g - 1) / g;",
     "double l2 = std::fabs(uavgn) * rho_log / g;", 2),
    ("matrix_l4_scale", "double l4 = std::fabs(uavgn) * p_avg;", "double l4 = std::fabs(uavgn) * rho_log;", 1),
    ("kg_gravity_weight", "double alpha = rho_avg / 2;", "double alpha = rho_avg;", 1),
    ("kg_momentum_mass_flux",
     "        double frho_e = u_avg[e] * rho_avg;\n"
     "        F[d][1 + e] = (d == e) ? u_avg[d] * frho_e + p_avg : u_avg[d] * frho_e;",
     "        double frho_e = avg(q1[1 + e], q2[1 + e]);\n"
     "        F[d][1 + e] = (d == e) ? u_avg[d] * frho_e + p_avg : u_avg[d] * frho_e;", 1),
    ("kg_energy_drop_pressure", "F[d][D + 1] = u_avg[d] * (rho_avg * e_avg + p_avg);",
     "F[d][D + 1] = u_avg[d] * (rho_avg * e_avg);", 1),
]


def _replace_nth(text, old, new, nth):
    pos = -1
    for _ in range(nth):
        pos = text.find(old, pos + 1)
        if pos < 0:
            raise AssertionError(f"mutation site not found: {old!r} (occurrence {nth})")
    return text[:pos] + new + text[pos + len(old):]


def _build(tmp, name, laws_text):
    d = os.path.join(tmp, name)
    os.makedirs(d)
    shutil.copy(os.path.join(ORACLE, "atum_oracle.cpp"), d)
    with open(os.path.join(d, "laws.hpp"), "w") as f:
        f.write(laws_text)
    so = os.path.join(d, "liboracle.so")
    subprocess.run(CXX + [os.path.join(d, "atum_oracle.cpp"), "-o", so], check=True, capture_output=True)
    return so


def _run_pins(so):
    env = dict(os.environ, ATUM_ORACLE_SO=so, OMP_NUM_THREADS="1")
    env.pop("PYTEST_XDIST_WORKER", None)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_flux_pins.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with open(os.path.join(ORACLE, "laws.hpp")) as f:
        laws = f.read()
    jobs = {"unmodified": laws}
    for name, old, new, nth in MUTANTS:
        jobs[name] = _replace_nth(laws, old, new, nth)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        futs = {name: ex.submit(_build, tmp, name, text) for name, text in jobs.items()}
        sos = {name: f.result() for name, f in futs.items()}
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        futs = {name: ex.submit(_run_pins, so) for name, so in sos.items()}
        return {name: f.result() for name, f in futs.items()}


def test_unmodified_copy_passes(builds):
    rc, out = builds["unmodified"]
    assert rc == 0, out


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_mutant_is_caught(builds, name):
    rc, out = builds[name]
    assert rc == 1, f"the flux checks did not fail on mutant {name}:\n{out}"
