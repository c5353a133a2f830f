// qvti_launch.h -- per-length launch table of the pencil kernels.
//
// The kernels are templates over the split G = L2<N1, N2, PW> of an axis
// length (qvti_pencils.h).  Each supported length gets one LenOps record of
// launch functions, defined in the build group of qvti_kernels.hip that the
// length belongs to (qvti_lengths.h); the groups compile in parallel.  Tile
// widths PW = 1, 2, 4 (256, 512, 1024 threads) are indexed 0, 1, 2; a width
// that a length does not support (short axes, LDS) has null entries.
#pragma once
#include <hip/hip_runtime.h>

#include "qvti_lengths.h"
#include "qvti_pencils.h"

namespace qvti {

struct LenOps {
    int n, n1, n2;
    int Prow;  // rows per workgroup of the row kernels (256 threads)
    int Rsz;   // rows per workgroup of the stress update (StressTile)
    int P[3];  // pencils per tile of the strided kernels at width index 0, 1, 2 (0: not available)
    int T[3];  // threads per workgroup at width index 0, 1, 2
    // strided-axis kernels, [..][pw index]
    void (*deriv[3][3])(dim3 grid, hipStream_t st, const StridedArgs& a);  // [MODE_*]
    void (*cfft[2][3])(dim3 grid, hipStream_t st, const CStridedArgs& a);  // [0 forward, 1 inverse]
    void (*xpass[2][3])(dim3 grid, hipStream_t st, const XPassArgs& a);    // [k-space mode off, on]
    // row (z) kernels, 256 threads
    void (*vel_z[2])(dim3 grid, hipStream_t st, const VelZArgs& a);  // [without, with extra point sources]
    void (*zrows)(dim3 grid, hipStream_t st, const ZRowsArgs& a);
    void (*stress_z[2])(dim3 grid, hipStream_t st, const StressZArgs& a);  // [isotropic, VTI]
    void (*deriv_rows)(dim3 grid, hipStream_t st, const RowDerivArgs& a);
};

// The record of axis length n with split variant 0 (default) or 1 (the
// alternative of QVTI_ALT_LENGTHS), or null when there is none.
const LenOps* len_ops(int n, int variant = 0);
// The supported lengths, space separated.
const char* supported_lengths();
// Widest available tile width index <= the requested width pw (1, 2 or 4).
inline int pw_index(const LenOps& o, int pw)
{
    const int want = pw >= 4 ? 2 : (pw >= 2 ? 1 : 0);
    for (int i = want; i > 0; --i)
        if (o.P[i] > 0) return i;
    return 0;
}

#define QVTI_DECLARE_LEN(n, a, b, g) extern const LenOps kLenOps_##n;
QVTI_LENGTHS(QVTI_DECLARE_LEN)
#undef QVTI_DECLARE_LEN
#define QVTI_DECLARE_ALT(n, a, b, g) extern const LenOps kLenOps_Alt_##n;
QVTI_ALT_LENGTHS(QVTI_DECLARE_ALT)
#undef QVTI_DECLARE_ALT

}  // namespace qvti
