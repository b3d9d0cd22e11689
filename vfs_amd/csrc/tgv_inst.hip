// One instantiation unit of the op kernels, compiled once per storage type and kernel
// family (the library build passes VFS_INST_STORAGE = 0 float, 1 __half, 2
// __hip_bfloat16, and VFS_INST_K = 1 for the tick kernel, 2, 4 or 8 for the K-iteration
// kernel, 0 for the whole-warp persistent kernel), so that the kernels compile in parallel.
// Defines launch_ops_<storage>_k<K>, which kernels.hip's launch_tick dispatches to (the
// K-iteration units also hold the blocked persistent op), and for K != 1
// persist_capacity_<storage>_k<K>.
#include "tgv_device.h"

#if !defined(VFS_INST_STORAGE) || !defined(VFS_INST_K)
#error "build with -DVFS_INST_STORAGE=<0|1|2> -DVFS_INST_K=<0|1|2|4|8>"
#endif

#if VFS_INST_STORAGE == 0
#define VFS_INST_T float
#define VFS_INST_SFX f32
#elif VFS_INST_STORAGE == 1
#define VFS_INST_T __half
#define VFS_INST_SFX f16
#else
#define VFS_INST_T __hip_bfloat16
#define VFS_INST_SFX bf16
#endif

#define VFS_CAT3(a, b, c) a##b##c
#define VFS_NAME(s, k) VFS_CAT3(launch_ops_, s, _k##k)
#define VFS_NAME_X(s, k) VFS_NAME(s, k)

namespace vfs {

void VFS_NAME_X(VFS_INST_SFX, VFS_INST_K)(const PairDesc *desc, const TickArgs &args, int shape, hipStream_t st)
{
#if VFS_INST_K == 1
    tick_storage<VFS_INST_T>(desc, args, shape, st);
#elif VFS_INST_K == 0
    iterp_storage<VFS_INST_T>(desc, args, shape, st);
#else
    if (args.e[0].kind == OP_ITERPK) iterkp_storage<VFS_INST_T, VFS_INST_K>(desc, args, shape, st);
    else iterk_storage<VFS_INST_T, VFS_INST_K>(desc, args, shape, st);
#endif
}

#define VFS_CAP_NAME(s, k) VFS_CAT3(persist_capacity_, s, _k##k)
#define VFS_CAP_NAME_X(s, k) VFS_CAP_NAME(s, k)
#if VFS_INST_K == 0
int VFS_CAP_NAME_X(VFS_INST_SFX, VFS_INST_K)(int shape) { return iterp_capacity<VFS_INST_T>(shape); }
#elif VFS_INST_K != 1
int VFS_CAP_NAME_X(VFS_INST_SFX, VFS_INST_K)(int shape) { return iterkp_capacity<VFS_INST_T, VFS_INST_K>(shape); }
#endif

}  // namespace vfs
