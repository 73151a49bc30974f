// Convergence bookkeeping of the mixed path's Ogita-Aishima refinement.
// Host-only C++ (no HIP): the decisions the refinement loop takes after each
// step, kept apart so that they can be exercised without a GPU.
#pragma once
#include <cmath>
#include <vector>

namespace gmc {

// The correction a problem reports after one step.  emax is the largest
// off-diagonal |E_ij| of the step as the device reduction left it (that
// reduction maps a NaN element to +inf, so a non-finite step reads +inf);
// npair is the number of close pairs the step found, of which only the first
// `cap` were recorded and rotated.  A problem with unrecorded pairs has
// corrections that emax does not show: it counts as not converged (+inf).
inline float oa_effective(float emax, unsigned npair, unsigned cap)
{
    if (npair > cap) return INFINITY;
    if (!(emax <= 3.4028235e38f)) return INFINITY;   // NaN or inf
    return emax;
}

// The convergence measure of a problem after one step: its largest
// correction, or its largest predicted next correction times OA_PRED_SCALE,
// whichever is larger.  The prediction of a separated pair is E_ij^2
// (max|l| / gap) weighted by how differently the solve uses the two
// eigenvalues (the device side computes it).  With the extend threshold 5e-4
// a problem counts as converged only when every pair's prediction is below
// 1e-5: a pair with a gap of 1e-5 of the spectrum, between a mode the
// integrand keeps and one it filters out, left 2e-6 Eh in W(alpha) after a
// last correction of 4e-5 (a 1494-pair system at alpha = 0.88).  A
// non-finite part reads +inf.
constexpr float OA_PRED_SCALE = 50.f;
inline float oa_measure(float emax, float pred)
{
    if (!(emax <= 3.4028235e38f) || !(pred <= 3.4028235e38f)) return INFINITY;
    return std::fmax(emax, pred * OA_PRED_SCALE);
}

enum class OaNext {
    Continue,   // run another step
    Stop,       // every problem converged (or the fixed count ran out with finite corrections)
    Resolve,    // stop; the problems in `left` are solved again in FP64
    Fail        // fixed step count, no FP64 re-solve, and some problem is not finite / over the pair cap
};

struct OaDecision {
    OaNext next = OaNext::Continue;
    std::vector<int> left;
};

// After step `it` (0-based) with per-problem corrections pe (oa_effective):
//   - steps 0 .. n_min-1 always run unless every correction is below tol;
//   - after them the refinement goes on while some correction is at least
//     `leave` (at most `extend`; by default equal to it) and below giveup,
//     up to n_max steps;
//   - with a re-solve available (adaptive mode), the problems whose
//     correction is not below extend when it stops go to `left` (a
//     correction in [leave, extend) counts as converged); it stops early
//     once every problem at or above leave is at or above giveup (not
//     converging);
//   - without one (a fixed step count), a problem left at +inf fails the call.
// Every comparison is written so that +inf and NaN count as not converged.
inline OaDecision oa_decide(const std::vector<float>& pe, int it, int n_min, int n_max, float tol, float extend,
                            float giveup, bool have_resolve, float leave = -1.f)
{
    if (!(leave >= 0.f) || leave > extend) leave = extend;
    OaDecision d;
    bool all_below_tol = true, all_below_leave = true, any_slow = false, any_bad = false;
    for (float e : pe) {
        if (!(e < tol)) all_below_tol = false;
        if (!(e < leave)) all_below_leave = false;
        if (e >= leave && e < giveup) any_slow = true;   // still converging
        if (!(e <= 3.4028235e38f)) any_bad = true;
    }
    const bool last = it + 1 >= n_max;
    if (it + 1 < n_min && !last) {
        d.next = all_below_tol ? OaNext::Stop : OaNext::Continue;
        return d;
    }
    if (!have_resolve) {
        if (!last) {
            d.next = all_below_tol ? OaNext::Stop : OaNext::Continue;
            return d;
        }
        if (any_bad) {
            for (int b = 0; b < (int)pe.size(); b++)
                if (!(pe[b] <= 3.4028235e38f)) d.left.push_back(b);
            d.next = OaNext::Fail;
        } else {
            d.next = OaNext::Stop;
        }
        return d;
    }
    if (all_below_leave) {
        d.next = OaNext::Stop;
        return d;
    }
    if (last || !any_slow) {
        for (int b = 0; b < (int)pe.size(); b++)
            if (!(pe[b] < extend)) d.left.push_back(b);
        d.next = d.left.empty() ? OaNext::Stop : OaNext::Resolve;
        return d;
    }
    d.next = OaNext::Continue;
    return d;
}

}  // namespace gmc
