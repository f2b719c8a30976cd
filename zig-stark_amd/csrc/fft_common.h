// fft_common.h -- pass descriptors shared by the FFT kernel files.
#pragma once
#include <algorithm>
#include <vector>

#include "internal.h"

namespace zs {

enum : uint32_t { IN_KSPACE = 0, IN_SHARED = 1, IN_NATURAL = 2, IN_AB = 3 };
enum : uint32_t { OUT_KSPACE = 0, OUT_NATURAL = 1, OUT_AB = 2 };

struct FftPass {
    const uint32_t *in;
    uint32_t *out;
    uint64_t in_cs, out_cs;  // column strides (elements)
    const uint32_t *tw;      // heap table of log size LT
    uint32_t L, LT, beta;    // transform log size, table log size, blowup log
    uint32_t s_lo, C, G_log, Q_log;
    uint32_t in_mode, out_mode;
    uint32_t scale;          // 0 = none
    uint32_t natural_tiles;  // tiles chosen so bitrev addresses are adjacent
    uint32_t n_cols;         // columns in the batch
    uint32_t ncb;            // columns one workgroup processes with the same staged twiddles
    uint32_t TS;             // LDS words from one tile to the next (register-group kernels)
};

struct PassRange {
    uint32_t s_lo, C;
};

// Split L stages into passes of at most cmax stages.  The top pass (the highest
// stages) has the same size for both directions, so an interpolation's last pass
// and an evaluation's first pass cover the same stage range and can be fused.
inline std::vector<PassRange> plan_passes(uint32_t L, bool inverse, uint32_t cmax) {
    uint32_t np = (L + cmax - 1) / cmax;
    std::vector<uint32_t> sizes(np, L / np);
    for (uint32_t i = 0; i < L % np; i++) sizes[i]++;
    std::vector<PassRange> out;
    if (inverse) {
        uint32_t s = 0;
        for (size_t i = np; i-- > 0;) { out.push_back({s, sizes[i]}); s += sizes[i]; }
    } else {
        uint32_t s = L;
        for (uint32_t c : sizes) { s -= c; out.push_back({s, c}); }
    }
    return out;
}

constexpr uint32_t PASS_CMAX = 11;
constexpr uint32_t PASS_ELEMS_LOG = 13;


// generic LDS pass (any C), and the register-group pass (C in 8..11)
int32_t run_pass(Ctx *ctx, FftPass p, bool inverse, uint32_t n_cols);
int32_t run_tile_pass(Ctx *ctx, FftPass p, bool inverse, uint32_t n_cols);
inline bool tile_pass_ok(uint32_t C) { return C >= 8 && C <= 11; }
// fused LDE middle: interpolation stages [s_lo, L) + scale + the 2^beta forward
// transforms' stages [L-1, s_lo]; in: natural (s_lo == 0) or k-space coefficients,
// out: natural LDE (s_lo == 0) or k-space big array
bool lde_fused_ok(uint32_t C, uint32_t beta, bool single);
// LDS words from one register-group tile to the next when `il` tiles share a copy instruction
uint32_t tile_stride(uint32_t C, uint32_t il);
int32_t run_lde_fused(Ctx *ctx, const uint32_t *d_in, uint64_t in_cs, uint32_t *d_out, uint64_t out_cs,
                      const uint32_t *tw_inv, const uint32_t *tw_fwd, uint32_t L, uint32_t beta, uint32_t s_lo,
                      uint32_t C, uint32_t n_cols);

}  // namespace zs
