// qvti_kernels.hip -- instantiates the pencil kernels of one build group of
// axis lengths (qvti_lengths.h) and defines their launch records
// (qvti_launch.h).  Compiled once per group with -DQVTI_GROUP=g.
#include "qvti_launch.h"

#ifndef QVTI_GROUP
#error "compile with -DQVTI_GROUP=<build group>"
#endif

namespace qvti {
namespace {

template <class G, int MODE>
void l_deriv(dim3 grid, hipStream_t st, const StridedArgs& a)
{
    hipLaunchKernelGGL((k2_deriv_strided<G, MODE>), grid, dim3(G::T), 0, st, a);
}
template <class G, int DIR>
void l_cfft(dim3 grid, hipStream_t st, const CStridedArgs& a)
{
    hipLaunchKernelGGL((k2_cfft_strided<G, DIR>), grid, dim3(G::T), 0, st, a);
}
template <class G, bool KS>
void l_xpass(dim3 grid, hipStream_t st, const XPassArgs& a)
{
    hipLaunchKernelGGL((k2_xpass<G, KS>), grid, dim3(G::T), 0, st, a);
}
template <class G, bool INJ>
void l_vel_z(dim3 grid, hipStream_t st, const VelZArgs& a)
{
    hipLaunchKernelGGL((k2_vel_z<G, INJ>), grid, dim3(kThreads), 0, st, a);
}
template <class G>
void l_zrows(dim3 grid, hipStream_t st, const ZRowsArgs& a)
{
    hipLaunchKernelGGL(k2_zrows<G>, grid, dim3(kThreads), 0, st, a);
}
template <class G, bool VTI>
void l_stress_z(dim3 grid, hipStream_t st, const StressZArgs& a)
{
    hipLaunchKernelGGL((k2_stress_z<G, VTI>), grid, dim3(kThreads), 0, st, a);
}
template <class G>
void l_deriv_rows(dim3 grid, hipStream_t st, const RowDerivArgs& a)
{
    hipLaunchKernelGGL(k2_deriv_rows<G>, grid, dim3(kThreads), 0, st, a);
}

// Tile width PW of the strided kernels exists for n >= 64 while the tile fits in LDS.
template <int A, int B, int PW>
constexpr bool pw_ok()
{
    if constexpr (PW == 1) return true;
    else return A * B >= 64 && L2<A, B, PW>::kFits;
}

using DerivFn = void (*)(dim3, hipStream_t, const StridedArgs&);
using CfftFn = void (*)(dim3, hipStream_t, const CStridedArgs&);
using XpassFn = void (*)(dim3, hipStream_t, const XPassArgs&);

template <int A, int B, int PW, int MODE>
DerivFn deriv_fn()
{
    if constexpr (pw_ok<A, B, PW>()) return &l_deriv<L2<A, B, PW>, MODE>;
    else return nullptr;
}
template <int A, int B, int PW, int DIR>
CfftFn cfft_fn()
{
    if constexpr (pw_ok<A, B, PW>()) return &l_cfft<L2<A, B, PW>, DIR>;
    else return nullptr;
}
template <int A, int B, int PW, bool KS>
XpassFn xpass_fn()
{
    if constexpr (pw_ok<A, B, PW>()) return &l_xpass<L2<A, B, PW>, KS>;
    else return nullptr;
}
template <int A, int B, int PW>
constexpr int pencils()
{
    if constexpr (pw_ok<A, B, PW>()) return L2<A, B, PW>::P;
    else return 0;
}

// Built at load time (dynamic initialisation), so that every launch function
// whose address the record holds is instantiated and emitted.
// tile widths PW of width index 0, 1, 2
constexpr int kW0 = 1, kW1 = 2, kW2 = 4;

template <int A, int B>
LenOps make_ops()
{
    using G = L2<A, B, 1>;
    return LenOps{
        A * B, A, B, G::P, StressTile<G>::R,
        {pencils<A, B, kW0>(), pencils<A, B, kW1>(), pencils<A, B, kW2>()},
        {L2<A, B, kW0>::T, L2<A, B, kW1>::T, L2<A, B, kW2>::T},
        {{deriv_fn<A, B, kW0, MODE_STORE>(), deriv_fn<A, B, kW1, MODE_STORE>(), deriv_fn<A, B, kW2, MODE_STORE>()},
         {deriv_fn<A, B, kW0, MODE_VEL_FIRST>(), deriv_fn<A, B, kW1, MODE_VEL_FIRST>(),
          deriv_fn<A, B, kW2, MODE_VEL_FIRST>()},
         {deriv_fn<A, B, kW0, MODE_VEL_ACC>(), deriv_fn<A, B, kW1, MODE_VEL_ACC>(), deriv_fn<A, B, kW2, MODE_VEL_ACC>()}},
        {{cfft_fn<A, B, kW0, -1>(), cfft_fn<A, B, kW1, -1>(), cfft_fn<A, B, kW2, -1>()},
         {cfft_fn<A, B, kW0, 1>(), cfft_fn<A, B, kW1, 1>(), cfft_fn<A, B, kW2, 1>()}},
        {{xpass_fn<A, B, kW0, false>(), xpass_fn<A, B, kW1, false>(), xpass_fn<A, B, kW2, false>()},
         {xpass_fn<A, B, kW0, true>(), xpass_fn<A, B, kW1, true>(), xpass_fn<A, B, kW2, true>()}},
        {&l_vel_z<G, false>, &l_vel_z<G, true>},
        &l_zrows<G>,
        {&l_stress_z<G, false>, &l_stress_z<G, true>},
        &l_deriv_rows<G>,
    };
}

}  // namespace

#define QVTI_DEFINE_IN_GROUP(n, a, b, g) QVTI_DEFINE_IN_GROUP_##g(n, a, b)
#define QVTI_DEFINE_LEN(n, a, b) const LenOps kLenOps_##n = make_ops<a, b>();
#define QVTI_SKIP_LEN(n, a, b)
#if QVTI_GROUP == 0
#define QVTI_DEFINE_IN_GROUP_0 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_0 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 1
#define QVTI_DEFINE_IN_GROUP_1 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_1 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 2
#define QVTI_DEFINE_IN_GROUP_2 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_2 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 3
#define QVTI_DEFINE_IN_GROUP_3 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_3 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 4
#define QVTI_DEFINE_IN_GROUP_4 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_4 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 5
#define QVTI_DEFINE_IN_GROUP_5 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_5 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 6
#define QVTI_DEFINE_IN_GROUP_6 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_6 QVTI_SKIP_LEN
#endif
#if QVTI_GROUP == 7
#define QVTI_DEFINE_IN_GROUP_7 QVTI_DEFINE_LEN
#else
#define QVTI_DEFINE_IN_GROUP_7 QVTI_SKIP_LEN
#endif

QVTI_LENGTHS(QVTI_DEFINE_IN_GROUP)

#define QVTI_ALT_IN_GROUP(n, a, b, g) QVTI_DEFINE_IN_GROUP_##g(Alt_##n, a, b)
QVTI_ALT_LENGTHS(QVTI_ALT_IN_GROUP)

}  // namespace qvti
