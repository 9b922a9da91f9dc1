// xdem_step_f64.hip -- FP64 instantiation of the contact and velocity kernels.
// Built with -ffp-contract=off so every multiply and add rounds separately, in the
// reference's operation order.
#include <algorithm>

#include "xdem_step.hpp"

namespace xdem {
namespace {

constexpr unsigned kBlock = 256;
inline dim3 grid_for(unsigned n) { return dim3((n + kBlock - 1) / kBlock); }

// pass 0: the lite pass over every sphere outside this step's full pass (two kernels in
// fused mode); pass 2: the triangle pass (meshes) and pass 3: the sphere pairs, both before
// pass 1, the rest of the full pass over the spheres k_touch queued. The full-pass kernels
// stride over the queue segments (the queue lengths are on the device) with A.full_blocks
// blocks, a multiple of kQueueSegs: a few blocks per CU keep several waves per SIMD for the
// neighbour gathers. The two passes touch disjoint spheres and may run concurrently.
inline dim3 full_grid(const StepArgs& A)
{
    const unsigned want = std::min(A.full_blocks, std::max(grid_for(A.np).x, kQueueSegs));
    return dim3(std::max(1u, want / kQueueSegs) * kQueueSegs);
}

template <bool MULTI, bool FUSE, int MODEL, bool GPU_LAW>
void collide_passes(const StepArgs& A, int pass, hipStream_t s)
{
    const dim3 g = grid_for(A.np), b(kBlock);
    if (pass == 1) {
        if (A.has_cyl) hipLaunchKernelGGL((k_collide<double, MULTI, FUSE, MODEL, GPU_LAW, true>), full_grid(A), b, 0, s, A);
        else hipLaunchKernelGGL((k_collide<double, MULTI, FUSE, MODEL, GPU_LAW, false>), full_grid(A), b, 0, s, A);
    } else if (pass == 3) {
        // the pair pass: the resident grid (every block in one round)
        hipLaunchKernelGGL((k_pairs<double, MULTI, MODEL, GPU_LAW>), dim3(A.resident_blocks), b, 0, s, A);
    } else if (pass == 2) {
        // the triangle pass (grid-strided over the mesh queue, whose length is on the device)
        hipLaunchKernelGGL((k_mesh_pass<double, MULTI, MODEL, GPU_LAW>), full_grid(A), b, 0, s, A);
    } else if (FUSE) {
        launch_lite_fused(A, MULTI, s);
    } else {
        hipLaunchKernelGGL((k_collide_lite<double, MULTI, false, MODEL, GPU_LAW>), g, b, 0, s, A);
    }
}

template <bool MULTI, bool FUSE>
void collide_models(const StepArgs& A, int model, bool gpu_law, int pass, hipStream_t s)
{
    if (model == 1) {
        if (gpu_law) collide_passes<MULTI, FUSE, 1, true>(A, pass, s);
        else collide_passes<MULTI, FUSE, 1, false>(A, pass, s);
    } else {
        if (gpu_law) collide_passes<MULTI, FUSE, 0, true>(A, pass, s);
        else collide_passes<MULTI, FUSE, 0, false>(A, pass, s);
    }
}

hipError_t collide_f64(const StepArgs& A, bool multi, bool fuse, int model, bool gpu_law, int pass,
                       hipStream_t s)
{
    if (A.np == 0) return hipSuccess;
    if (multi) {
        if (fuse) collide_models<true, true>(A, model, gpu_law, pass, s);
        else collide_models<true, false>(A, model, gpu_law, pass, s);
    } else {
        if (fuse) collide_models<false, true>(A, model, gpu_law, pass, s);
        else collide_models<false, false>(A, model, gpu_law, pass, s);
    }
    return hipGetLastError();
}

}  // namespace

CollideLaunch collide_launch_f64() { return collide_f64; }

void launch_lite_fused(const StepArgs& A, bool multi, hipStream_t s, int which)
{
    const dim3 g = grid_for(A.np), b(kBlock);
    if (which != 2) {
        if (multi) hipLaunchKernelGGL(k_lite_move<true>, g, b, 0, s, A);
        else hipLaunchKernelGGL(k_lite_move<false>, g, b, 0, s, A);
    }
    if (which != 1) {
        if (multi) hipLaunchKernelGGL(k_lite_turn<true>, g, b, 0, s, A);
        else hipLaunchKernelGGL(k_lite_turn<false>, g, b, 0, s, A);
    }
}

hipError_t launch_finish(const StepArgs& A, bool multi, hipStream_t s)
{
    if (A.np == 0) return hipSuccess;
    if (multi) hipLaunchKernelGGL(k_finish<true>, grid_for(A.np), dim3(kBlock), 0, s, A);
    else hipLaunchKernelGGL(k_finish<false>, grid_for(A.np), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

}  // namespace xdem
