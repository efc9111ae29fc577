// Fused time-step kernels instantiated for float storage.
#include "step_impl.h"

namespace mkg {

#define MKG_CASE(FQ, PQ, TH, SO)                                          \
    if (k.fq == FQ && k.pq == PQ && k.th == TH && k.so == SO) {           \
        run_step<float, FQ, PQ, (bool)TH, (bool)SO>(odd, a, s);              \
        return 0;                                                         \
    }

int launch_step_f32(const Combo& k, int odd, const StepArgs& a, hipStream_t s) {
    MKG_COMBOS(MKG_CASE)
    return -1;
}

#define MKG_NEE_CASE(FQ, PQ, TH, SO)                                      \
    if (k.fq == FQ && k.pq == PQ && k.th == TH && k.so == SO) {           \
        run_nee<float, FQ, PQ, (bool)TH>(a, s);                              \
        return 0;                                                         \
    }

int launch_nee_f32(const Combo& k, const StepArgs& a, hipStream_t s) {
    MKG_COMBOS(MKG_NEE_CASE)
    return -1;
}

int launch_phase_moments_f32(int pq, int odd, const StepArgs& a, void* out, hipStream_t s) {
    if (pq == 19) {
        run_phase_moments<float, 19>(odd, a, out, s);
        return 0;
    }
    if (pq == 7) {
        run_phase_moments<float, 7>(odd, a, out, s);
        return 0;
    }
    return -1;
}

}  // namespace mkg
