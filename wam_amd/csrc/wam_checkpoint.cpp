// wam_checkpoint.cpp -- checkpointing schedules for the adjoint of a time
// window of model steps.
//
// Reversing NSTEPS steps needs the state at the start of every step in
// reverse order.  With NSNAP checkpoint slots besides the one that holds the
// window's initial state, the schedule that recomputes the fewest forward
// steps splits the window recursively: advance m steps from a held state,
// store a checkpoint there, reverse the right part with one slot fewer, then
// restore and reverse the left part with the slots the right part returned.
// The minimal number of recomputed steps obeys
//     c(1, s) = 0,  c(l, 0) = l (l - 1) / 2,
//     c(l, s) = min over 1 <= m < l of  m + c(l - m, s - 1) + c(m, s),
// whose optimum is the binomial (revolve) schedule; it is tabulated here by
// dynamic programming and the schedule is emitted as actions for the caller
// to execute.  (The reference has no adjoint and so no counterpart.)
#include <algorithm>
#include <cstdint>
#include <vector>
#include "../../include/wam_gpu.h"

namespace {

struct Revolve {
    int L = 0, S = 0;
    std::vector<int64_t> cost;      // [S + 1][L + 1]
    std::vector<int32_t> split;     // [S + 1][L + 1]
    std::vector<int32_t> free_slots;
    int32_t *out = nullptr;
    int64_t cap = 0, n = 0;

    int64_t &c(int s, int l) { return cost[(size_t)s * (L + 1) + l]; }
    int32_t &m_of(int s, int l) { return split[(size_t)s * (L + 1) + l]; }

    void table()
    {
        cost.assign((size_t)(S + 1) * (L + 1), 0);
        split.assign((size_t)(S + 1) * (L + 1), 0);
        for (int l = 1; l <= L; l++) c(0, l) = (int64_t)l * (l - 1) / 2;
        for (int s = 1; s <= S; s++)
            for (int l = 2; l <= L; l++) {
                if (s >= l - 1) {           // a slot for every state: store them all on the way
                    c(s, l) = l - 1;
                    m_of(s, l) = 1;
                    continue;
                }
                int64_t best = -1;
                int bm = 1;
                for (int m = 1; m < l; m++) {
                    const int64_t v = m + c(s - 1, l - m) + c(s, m);
                    if (best < 0 || v < best) { best = v; bm = m; }
                }
                c(s, l) = best;
                m_of(s, l) = bm;
            }
    }

    void emit(int op, int a, int b)
    {
        if (out && n < cap) {
            out[3 * n] = op;
            out[3 * n + 1] = a;
            out[3 * n + 2] = b;
        }
        n++;
    }

    // reverse steps a .. a + l - 1; the current state is step a, held in slot
    void rev(int a, int l, int s, int slot)
    {
        if (l == 1) {
            emit(WAM_CK_ADJOINT, a, 0);
            return;
        }
        if (s == 0) {
            for (int i = l - 1; i >= 0; i--) {
                if (i != l - 1) emit(WAM_CK_RESTORE, a, slot);
                if (i > 0) emit(WAM_CK_ADVANCE, a, a + i);
                emit(WAM_CK_ADJOINT, a + i, 0);
            }
            return;
        }
        const int m = m_of(s, l);
        emit(WAM_CK_ADVANCE, a, a + m);
        const int32_t fresh = free_slots.back();
        free_slots.pop_back();
        emit(WAM_CK_STORE, a + m, fresh);
        rev(a + m, l - m, s - 1, fresh);
        free_slots.push_back(fresh);
        emit(WAM_CK_RESTORE, a, slot);
        rev(a, m, s, slot);
    }
};

// the table costs about L*L*S/2 steps of its inner loop: refuse what would
// take more than a few seconds
bool too_long(int L, int S) { return (double)L * L * S / 2 > 4e9; }

}  // namespace

extern "C" {

int64_t wam_checkpoint_schedule(int32_t nsteps, int32_t nsnap, int32_t *actions, int64_t cap)
{
    if (nsteps < 1 || nsteps > 20000 || nsnap < 0 || cap < 0) return -WAM_E_ARG;
    Revolve r;
    r.L = nsteps;
    r.S = std::min(nsnap, nsteps - 1);
    if (too_long(r.L, r.S)) return -WAM_E_ARG;
    r.table();
    for (int32_t i = r.S; i >= 1; i--) r.free_slots.push_back(i);
    r.out = actions;
    r.cap = cap;
    r.emit(WAM_CK_STORE, 0, 0);
    r.rev(0, nsteps, r.S, 0);
    return r.n;
}

int64_t wam_checkpoint_recomputed_steps(int32_t nsteps, int32_t nsnap)
{
    if (nsteps < 1 || nsteps > 20000 || nsnap < 0) return -WAM_E_ARG;
    Revolve r;
    r.L = nsteps;
    r.S = std::min(nsnap, nsteps - 1);
    if (too_long(r.L, r.S)) return -WAM_E_ARG;
    r.table();
    return r.c(r.S, nsteps);
}

}  // extern "C"
