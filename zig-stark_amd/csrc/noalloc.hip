// noalloc.hip -- zs_m31_lde_commit_dev_noalloc: the LDE + row-leaf commitment of
// zs_m31_lde_commit_dev for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Everything the call needs beyond its inputs and outputs lives in one file-scope
// device array per GPU, reserved for the life of the process:
//   [0, 2^23) words        forward twiddle tables of canonicHalf(L), L = 3..22, the
//                          table of log size L at [2^L, 2^(L+1)) (heap layout);
//   [2^23, 2^24) words     the inverse tables, same placement;
//   [2^24, +48 Mi) words   the LDE's interleaved coefficients and, for transforms of
//                          more than one pass, its intermediate evaluations.
// The tables are written once per device at set-up and only read afterwards; the
// buffer region keeps whatever the last call left (every word the LDE reads there
// is written earlier in the same call, so nothing needs clearing).  Calls that
// overlap on one GPU (several streams or threads) share the buffer region and race.
#include <mutex>

#include "fft_common.h"

namespace zs {

constexpr uint32_t kNaTwLogMax = 22;                                 // largest table: 2^22 points
constexpr size_t kNaTwWords = size_t(2) << (kNaTwLogMax + 1);        // forward then inverse tables
constexpr size_t kNaBufWords = size_t(48) << 20;                     // LDE buffers, 192 MiB
constexpr size_t kNaWords = kNaTwWords + kNaBufWords;                // 256 MiB in all
constexpr int kNaMaxDevices = 64;

}  // namespace zs

static __device__ __align__(256) uint32_t scratch[zs::kNaWords];

namespace zs {
namespace {

// Per device: the array's device address (from hipGetSymbolAddress; the host-side
// name `scratch` is not a device address) and whether set-up succeeded.
struct NoAllocDevice {
    std::once_flag once;
    uint32_t *ptr = nullptr;
    int32_t status = ZS_ERR_DEVICE;
    std::string error = "allocation-free LDE: device not set up";
};
NoAllocDevice g_dev[kNaMaxDevices];

uint32_t *tw_fwd(uint32_t *base, uint32_t L) { return base + (size_t(1) << L); }
uint32_t *tw_inv(uint32_t *base, uint32_t L) { return base + (kNaTwWords >> 1) + (size_t(1) << L); }

void setup_device(NoAllocDevice &d) {
    uint32_t *p = nullptr;
    hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        d.error = std::string("allocation-free LDE: hipGetSymbolAddress: ") + hipGetErrorString(e);
        return;
    }
    hipStream_t s = nullptr;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) {
        d.error = std::string("allocation-free LDE: stream: ") + hipGetErrorString(e);
        return;
    }
    for (uint32_t L = 3; L <= kNaTwLogMax && e == hipSuccess; L++) {
        zs_circle_coset hc;
        zs_circle_coset_canonic_half(L, &hc);
        e = launch_twiddle_gen(s, hc, tw_fwd(p, L), tw_inv(p, L));
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        d.error = std::string("allocation-free LDE: twiddle tables: ") + hipGetErrorString(e);
        return;
    }
    d.ptr = p;
    d.status = ZS_OK;
}

}  // namespace

// Called by zs_hip_init with `device` current: once per device, the array's address
// and the twiddle tables.  A failure here only disables the allocation-free entry point.
void noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return;
    NoAllocDevice &d = g_dev[device];
    std::call_once(d.once, [&] { setup_device(d); });
}

}  // namespace zs

using namespace zs;

extern "C" int32_t zs_m31_lde_commit_dev_noalloc(zs_hip_ctx *h, const uint32_t *d_cols, uint32_t n_cols,
                                                 uint32_t log_n, uint32_t log_blowup, uint32_t *d_lde,
                                                 uint8_t *d_nodes) {
    if (!h || !d_cols || !d_lde || !d_nodes || n_cols == 0 || log_n < 1 || log_sum_over(log_n, log_blowup, 28))
        return ZS_ERR_INPUT;
    Ctx *ctx = &h->c;
    const uint32_t LT = log_n + log_blowup;
    if (LT > kNaTwLogMax) {
        ctx->set_error("allocation-free LDE: log_n + log_blowup = " + std::to_string(LT) + " exceeds " +
                       std::to_string(kNaTwLogMax));
        return ZS_ERR_INPUT;
    }
    size_t coef_words;
    const size_t big_words = lde_scratch_words(n_cols, log_n, log_blowup, &coef_words);
    if (coef_words + big_words > kNaBufWords) {
        ctx->set_error("allocation-free LDE: " + std::to_string((coef_words + big_words) * 4) +
                       " bytes of LDE buffers needed, the fixed array holds " + std::to_string(kNaBufWords * 4));
        return ZS_ERR_INPUT;
    }
    if (ctx->device < 0 || ctx->device >= kNaMaxDevices) return ZS_ERR_DEVICE;
    NoAllocDevice &d = g_dev[ctx->device];
    if (d.status != ZS_OK) {
        ctx->set_error(d.error);
        return d.status;
    }
    ctx->lde_tiles = lde_tiles_for(ctx->device, n_cols, log_n, log_blowup);
    uint32_t *buf = d.ptr + kNaTwWords;
    const LdeTables tabs{log_n >= 3 ? tw_inv(d.ptr, log_n) : nullptr, LT >= 3 ? tw_fwd(d.ptr, LT) : nullptr};
    const size_t n = size_t(1) << log_n, N = size_t(1) << LT;
    ZS_TRY(launch_lde(ctx, d_cols, n, n_cols, log_n, log_blowup, d_lde, N, buf, big_words ? buf + coef_words : nullptr,
                      &tabs));
    return launch_row_commit(ctx, d_lde, N, n_cols, LT, reinterpret_cast<uint32_t *>(d_nodes));
}
