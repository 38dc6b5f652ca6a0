// tran_lds.hip — the LDS-resident batched transient kernel (k_tran_lds).
//
// A persistent grid of one-wavefront workgroups walks a work queue of
// instances (one size class per launch, largest first).  A workgroup copies
// its instance's plan image into LDS — skipped when the previous instance
// used the same circuit — and runs the whole transient there (lds_core.h).
#include "lds_core.h"

namespace cedar {

namespace {

using namespace lds;

template <int NR>
__global__ __launch_bounds__(W) void k_tran_lds(LdsArgs A) {
    extern __shared__ double lds_mem[];
    uint16_t* blob = reinterpret_cast<uint16_t*>(lds_mem + A.dbl_max);
    int cur = -1;
    for (;;) {
        int qi = 0;
        if (threadIdx.x == 0) qi = atomicAdd(A.queue, 1);
        qi = __builtin_amdgcn_readfirstlane(__shfl(qi, 0, W));
        if (qi >= A.n_inst) break;
        const int inst = __builtin_amdgcn_readfirstlane(A.order[qi]);
        const int cid = __builtin_amdgcn_readfirstlane(A.inst_cid[inst]);
        const CPlan c = load_header(A.cplans + cid);
        if (cid != cur) {
            copy_image(c, blob);
            cur = cid;
            __syncthreads();
        }
        const View v = make_view(lds_mem, blob, c);
        NoRec rec;
        run_lds<NR>(rec, c, A.plans + cid, A.params + A.p_off[inst],
                    A.u0 ? A.u0 + A.u_off[inst] : nullptr,
                    A.wave ? A.wave + A.w_off[inst] : nullptr, A.u_final + A.u_off[inst],
                    A.stats + inst, A.o, v);
    }
}

}  // namespace

size_t lds_kernel_bytes(int32_t dbl_max, int32_t words_max) {
    return 8 * (size_t)dbl_max + 2 * (size_t)words_max;
}

int lds_max_blocks(int nr, size_t lds) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (nr) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_lds<1>, W, lds); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_lds<2>, W, lds); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_lds<4>, W, lds); break;
        default: break;
    }
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_tran_lds(int nr, const LdsArgs& a, int nblocks, size_t lds, hipStream_t s) {
    switch (nr) {
        case 1: hipLaunchKernelGGL(k_tran_lds<1>, dim3(nblocks), dim3(W), lds, s, a); break;
        case 2: hipLaunchKernelGGL(k_tran_lds<2>, dim3(nblocks), dim3(W), lds, s, a); break;
        case 4: hipLaunchKernelGGL(k_tran_lds<4>, dim3(nblocks), dim3(W), lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cedar
