/*
 * libvfs_dllhook: the reference's DLL entry points (dllhook/vfs.h:38-45,
 * dllhook/vfs.cpp:3-57) over libvfs, for consumers that load the solver through
 * Create / Init / SolveStereo (a plugin host calling into a shared library).
 * Signatures match the reference; the StereoTgv* handle is an opaque pointer here.
 */
#ifndef VFS_DLLHOOK_H
#define VFS_DLLHOOK_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vfs_dll_state StereoTgvHandle;

/* dllhook/vfs.cpp:3-6: a new solver handle (NULL on failure). */
StereoTgvHandle *Create(void);

/* dllhook/vfs.cpp:8-30: initialize, store focal and baseline, load the fisheye mask
 * (8-bit grayscale PNG scaled by 1/255, or binary PGM) and the trajectory and
 * calibration fields (Middlebury .flo).  Failures are reported by VfsDllLastError. */
void Init(StereoTgvHandle *vfs, int stereoWidth, int stereoHeight, float beta, float gamma, float alpha0,
          float alpha1, float timeStepLambda, float lambda, int nLevel, float fScale, int nWarpIters,
          int nSolverIters, float focal, float baseline, const char *fisheyeMaskFilename,
          const char *translationVectorFilename, const char *calibrationVectorFilename);

/* dllhook/vfs.cpp:32-57: equalizeHist of both 8-bit frames, copy in, masked solve.
 * The reference leaves `output` unwritten (its copy-out is commented out, vfs.cpp:51-56);
 * here it receives an 8-bit depth map: round(255 * min(Z, maxDepth) / maxDepth) with
 * Z = focal * baseline / u for the 1-D disparity u > 0 inside the mask, 0 elsewhere. */
void SolveStereo(StereoTgvHandle *vfs, unsigned char *fs1, unsigned char *fs2, unsigned char *output,
                 int stereoWidth, int stereoHeight, float maxDepth);

/* No counterpart in the reference (whose entry points exit the process on errors):
 * "" after a successful Init / SolveStereo, else the reason of the last failure. */
const char *VfsDllLastError(const StereoTgvHandle *vfs);

/* No counterpart in the reference (whose handles are never released): frees a handle. */
void VfsDllDestroy(StereoTgvHandle *vfs);

#ifdef __cplusplus
}
#endif

#endif
