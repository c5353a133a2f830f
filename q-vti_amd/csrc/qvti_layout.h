// qvti_layout.h -- HBM layout constants shared by the setup kernels and the
// pencil kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace qvti {

// Blocked field groups.  A group of nf per-cell fields is stored as 64-cell
// blocks with the fields interleaved: element (f, ip) sits at
// ((ip >> 6) * nf + f) * 64 + (ip & 63).  A thread reaches every field of its
// cell from one address with immediate offsets f * 256 bytes, and the fields
// of 64 neighbouring cells form one contiguous run of nf * 256 bytes.
// A group of N cells takes blk_len(N) floats per field: N rounded up to whole
// blocks (N need not be a multiple of 64: axis lengths such as 18 or 90 are
// supported).
constexpr int kBlk = 64;
__host__ __device__ __forceinline__ long long blk_len(long long n) { return (n + kBlk - 1) / kBlk * kBlk; }
__host__ __device__ __forceinline__ long long blk_index(long long ip, int nf, int f)
{
    return ((ip >> 6) * nf + f) * kBlk + (ip & 63);
}
// coefficient groups: G1 = {P 0-4, S 5-9, S-bar 10-14}, G2 = {H 0-4, D 5-9} (VTI)
constexpr int kG1 = 15, kG2 = 10;

}  // namespace qvti
