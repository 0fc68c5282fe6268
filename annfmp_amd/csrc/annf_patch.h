// Device helpers shared by the kernels that read p x p RGB patches straight
// from an image (the leaf k-NN and the fused encoder).
//
// A patch of an image is `rows` runs of `rowlen` bytes, `rstride` apart;
// kernels keep it as rows of rw = ceil(rowlen / 4) dwords with the tail of
// each row zero.  Staging is split in two steps so that every global load of
// a block is in flight before the first wait: fetch_rows issues the loads of
// this lane's rows (q, q + step, ...) as aligned windows of LW dwords into
// registers; for_row_words then aligns them into whole row dwords.
#pragma once

#include "annf_internal.h"

namespace annf {

// byte offset of window i (image mode: ncols windows per image row); window
// indices are below 2^31, so the row split is a 32-bit division
__device__ __forceinline__ long long patch_origin(const PatchGeom &g, long long i) {
    const unsigned ui = (unsigned)i, nc = (unsigned)g.ncols;
    const unsigned y = ui / nc, x = ui - y * nc;
    return ((long long)y * g.width + x) * 3;
}

// Rows per lane (RL, with a quad of lanes per patch) and dwords per row
// window (LW >= rw + 1) for the padded patch lengths DK of the kernels:
// 176 holds p <= 7, 256 p <= 9, 400 p <= 11, 512 p = 12.
template <int DK>
struct StageShape {
    static constexpr int RL = DK <= 176 ? 2 : 3;
    static constexpr int LW = DK <= 176 ? 7 : (DK <= 256 ? 8 : 10);
};

// Loads rows q, q + step, ... (RL of them at most) of the patch at byte
// offset o as LW-dword windows starting at the dword below each row; lanes
// with !valid, and rows past g.rows, load nothing and read as zero.
template <int RL, int LW>
__device__ __forceinline__ void fetch_rows(const PatchGeom &g, long long o, bool valid, int q, int step,
                                           uint32_t (&w)[RL][LW]) {
#pragma unroll
    for (int i = 0; i < RL; ++i) {
        const int r = q + step * i;
        const long long al = (o + (long long)r * g.rstride) & ~3LL;
#pragma unroll
        for (int k = 0; k < LW; ++k) w[i][k] = 0u;
        if (valid && r < g.rows) {
            if (al + 4LL * LW <= g.nbytes) {
                const uint32_t *src = reinterpret_cast<const uint32_t *>(g.base + al);
#pragma unroll
                for (int k = 0; k < LW; ++k) w[i][k] = src[k];
            } else {
                // the last rows of the image: bytes past its end read as zero
#pragma unroll
                for (int k = 0; k < LW; ++k) {
                    uint32_t v = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const long long ob = al + 4 * k + b;
                        if (ob < g.nbytes) v |= (uint32_t)g.base[ob] << (8 * b);
                    }
                    w[i][k] = v;
                }
            }
        }
    }
}

// fn(r, k, v) for dword k < rw of each fetched row r (tail bytes masked);
// needs rw <= LW - 1.
template <int RL, int LW, typename Fn>
__device__ __forceinline__ void for_row_words(const PatchGeom &g, long long o, const uint32_t (&w)[RL][LW], int q,
                                              int step, int rw, uint32_t tmask, Fn &&fn) {
#pragma unroll
    for (int i = 0; i < RL; ++i) {
        const int r = q + step * i;
        if (r < g.rows) {
            const int sh = (int)((o + (long long)r * g.rstride) & 3);
#pragma unroll
            for (int k = 0; k < LW - 1; ++k) {
                if (k < rw) {
                    uint32_t v = __builtin_amdgcn_alignbyte(w[i][k + 1], w[i][k], sh);
                    if (k == rw - 1) v &= tmask;
                    fn(r, k, v);
                }
            }
        }
    }
}

}  // namespace annf
