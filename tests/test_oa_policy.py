"""Host logic of the mixed path's refinement loop (gammcor_amd/csrc/oa_policy.hpp),
compiled with g++ and run on the CPU: which problems count as converged, when
the loop stops, which problems go to the FP64 re-solve and when a fixed step
count fails the call.  The cases are the ones a silent slip would hide: close
pairs beyond the recorded cap, NaN corrections, and +inf.
"""
import os
import subprocess
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "gammcor_amd", "csrc")

DRIVER = textwrap.dedent(r"""
    #include <cstdio>
    #include <cmath>
    #include "oa_policy.hpp"
    using namespace gmc;
    static const char* nm(OaNext n) {
        switch (n) { case OaNext::Continue: return "continue"; case OaNext::Stop: return "stop";
                     case OaNext::Resolve: return "resolve"; default: return "fail"; }
    }
    static void show(const char* tag, std::vector<float> pe, int it, int nmin, int nmax, bool res,
                     float leave = -1.f) {
        OaDecision d = oa_decide(pe, it, nmin, nmax, 1e-8f, 5e-4f, 5e-2f, res, leave);
        std::printf("%s %s", tag, nm(d.next));
        for (int b : d.left) std::printf(" %d", b);
        std::printf("\n");
    }
    int main() {
        // oa_effective: over the cap, NaN and inf all read +inf
        std::printf("eff %g %g %g %g\n", oa_effective(1e-3f, 10, 4096), oa_effective(1e-3f, 4097, 4096),
                    oa_effective(NAN, 0, 4096), oa_effective(INFINITY, 0, 4096));
        std::printf("eff0 %g\n", oa_effective(2e-9f, 1, 0));
        // oa_measure: the prediction counts 50-fold; non-finite parts read +inf
        std::printf("meas %g %g %g %g\n", oa_measure(4e-5f, 1.8e-4f), oa_measure(2.8e-4f, 1e-6f),
                    oa_measure(1e-3f, NAN), oa_measure(NAN, 0.f));
        // adaptive mode: n_min 3, n_max 6, re-solve available
        show("early_continue", {1e-3f, 1e-2f}, 0, 3, 6, true);
        show("early_stop", {1e-9f, 1e-10f}, 0, 3, 6, true);
        show("early_inf", {1e-9f, INFINITY}, 0, 3, 6, true);
        show("converged", {1e-4f, 2e-4f}, 2, 3, 6, true);
        show("extend", {1e-4f, 1e-3f}, 2, 3, 6, true);
        show("giveup", {1e-4f, 0.12f, INFINITY}, 2, 3, 6, true);
        show("last", {1e-4f, 1e-3f, 2e-5f}, 5, 3, 6, true);
        show("overflow_last", {1e-4f, oa_effective(1e-6f, 5000, 4096)}, 5, 3, 6, true);
        show("nan_mid", {1e-4f, oa_effective(NAN, 0, 4096), 1e-3f}, 3, 3, 6, true);
        // leave threshold 1e-6 below extend: corrections in [1e-6, 5e-4) keep
        // refining until n_max, then count as converged
        show("leave_mid", {2e-5f, 1e-7f}, 2, 3, 6, true, 1e-6f);
        show("leave_done", {5e-7f, 1e-7f}, 2, 3, 6, true, 1e-6f);
        show("leave_last", {2e-5f, 1e-3f, 1e-7f}, 5, 3, 6, true, 1e-6f);
        show("leave_last_ok", {2e-5f, 1e-7f}, 5, 3, 6, true, 1e-6f);
        show("leave_giveup", {0.12f, 1e-7f}, 3, 3, 6, true, 1e-6f);
        // fixed count (no re-solve): n_min == n_max == 2
        show("fixed_mid", {1e-3f, INFINITY}, 0, 2, 2, false);
        show("fixed_end", {1e-3f, 1e-2f}, 1, 2, 2, false);
        show("fixed_end_bad", {1e-3f, INFINITY, NAN}, 1, 2, 2, false);
        return 0;
    }
""")


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    d = tmp_path_factory.mktemp("oa_policy")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", HDR, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        t = line.split()
        res[t[0]] = t[1:]
    return res


def test_effective_correction(results):
    assert results["eff"] == ["0.001", "inf", "inf", "inf"]
    # a cap of 0 records no pair: any close pair makes the problem unconverged
    assert results["eff0"] == ["inf"]


def test_measure_counts_the_predicted_error(results):
    """A last correction of 4e-5 with a pair predicted at 1.8e-4 (a gap of
    1e-5 of the spectrum) is not converged at the 5e-4 threshold; 2.8e-4
    with well separated pairs is; non-finite parts are +inf."""
    m = [float(x) for x in results["meas"]]
    assert m[0] >= 5e-4 and m[1] < 5e-4 and m[1] == pytest.approx(2.8e-4, rel=1e-6)
    assert m[2] == float("inf") and m[3] == float("inf")


def test_adaptive_mode(results):
    assert results["early_continue"] == ["continue"]
    assert results["early_stop"] == ["stop"]
    assert results["early_inf"] == ["continue"]
    assert results["converged"] == ["stop"]
    assert results["extend"] == ["continue"]
    # 0.12 and +inf are both past the give-up level: stop and re-solve them
    assert results["giveup"] == ["resolve", "1", "2"]
    assert results["last"] == ["resolve", "1"]


def test_leave_threshold(results):
    """Below extend but above leave: the refinement continues (only those
    problems, in the path), and at n_max such a problem counts as converged;
    only a correction at or above extend goes to FP64."""
    assert results["leave_mid"] == ["continue"]
    assert results["leave_done"] == ["stop"]
    assert results["leave_last"] == ["resolve", "1"]
    assert results["leave_last_ok"] == ["stop"]
    assert results["leave_giveup"] == ["resolve", "0"]


def test_overflow_and_nan_go_to_fp64(results):
    # over the pair cap: small emax, but its unrecorded rotations are missing
    assert results["overflow_last"] == ["resolve", "1"]
    # a NaN problem is never taken as converged; the slow one keeps going
    assert results["nan_mid"] == ["continue"]


def test_fixed_count_fails_on_bad_problem(results):
    assert results["fixed_mid"] == ["continue"]
    assert results["fixed_end"] == ["stop"]
    assert results["fixed_end_bad"] == ["fail", "1", "2"]
