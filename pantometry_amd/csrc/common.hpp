// Shared plumbing of the gfx950 library: error reporting, device buffers, the violation record
// and the fixed-order reductions every audit uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>

#include "../../include/pantometry_amd.h"

namespace pm {

// The message behind the last PM_ERR_* status on this thread.
void set_error(const char *fmt, ...);

// A HIP call that must succeed; on failure records the message and returns PM_ERR_DEVICE from
// the enclosing function.
#define PM_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t pm_e_ = (call);                                                           \
        if (pm_e_ != hipSuccess) {                                                           \
            ::pm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(pm_e_), __FILE__, \
                            __LINE__);                                                       \
            return PM_ERR_DEVICE;                                                            \
        }                                                                                    \
    } while (0)

#define PM_CHECK(status)                  \
    do {                                  \
        int pm_s_ = (status);             \
        if (pm_s_ != PM_OK) return pm_s_; \
    } while (0)

// Confirms a gfx950 device is current; PM_ERR_DEVICE (with the reason) otherwise.  There is no
// CPU fallback anywhere in this library.
int require_device();

// Owning device allocation.
// Raw buffer accesses at a uniform base address and a 32-bit byte offset below 2^32 - 8.  The
// resource's last word sets a 32-bit data format (untyped loads and stores need a non-zero one);
// the record count admits every 32-bit offset.  Aux 2 marks a store non-temporal.
constexpr int BUF_WORD3 = 0x00020000;
__device__ __forceinline__ double buffer_load_f64(const double *base, uint32_t off) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, -1, BUF_WORD3);
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void buffer_store_f64_nt(double v, double *base, uint32_t off) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, -1, BUF_WORD3);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, (int)off, 0, 2);
}
// The same through a resource built once: the lane's offset `voff` plus a uniform offset `soff`
// (an SGPR operand of the instruction), their sum below 2^32 - 8.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const double *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, -1, BUF_WORD3);
}
__device__ __forceinline__ double buffer_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void buffer_store_f64_nt(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, (int)voff, (int)soff, 2);
}

class DevBuf {
  public:
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr, o.bytes_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_, bytes_ = o.bytes_;
            o.p_ = nullptr, o.bytes_ = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t bytes) {
        release();
        if (bytes == 0) return PM_OK;
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
            return PM_ERR_DEVICE;
        }
        p_ = p, bytes_ = bytes;
        return PM_OK;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr, bytes_ = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p_); }
    double *d() const { return static_cast<double *>(p_); }
    size_t bytes() const { return bytes_; }
    explicit operator bool() const { return p_ != nullptr; }

  private:
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

void fill_violation(pm_violation *v, const char *quantity, const char *site, double before,
                    double after, double scale, double tolerance, int64_t step);

// Violation::at (conserved.rs:148-157): a refusal that is not a before/after comparison.
inline void violation_at(pm_violation *v, const char *site, const char *quantity, double detail,
                         int64_t step) {
    fill_violation(v, quantity, site, detail, detail, std::fabs(detail), 0.0, step);
}

// The device-side state of an audited run: the previous ledger total and scale, a halt flag
// that freezes the state after a breach, and the breach record.  Lives in device memory.
struct AuditState {
    double prev_total;
    double prev_scale;
    double tolerance;
    int halted;
    int pad;
    int64_t breach_step;
    double before, after, scale;
};

// ---- device helpers ------------------------------------------------------------------------

// A workgroup barrier for data exchanged through LDS only: waits for this wave's LDS operations
// (lgkmcnt) and meets the other waves, but leaves global loads in flight.  __syncthreads()
// carries a workgroup-scope release fence, which also drains every outstanding global load
// (vmcnt(0)) and so cancels any prefetch across the barrier.  The memory clobber keeps the
// compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The harmonic face mean of solid.rs:706-713, `2.0 * a * b / sum` guarded at zero.  IEEE
// division (the correctly rounded gfx950 sequence); contraction is off for the whole library.
__device__ __forceinline__ double series(double a, double b) {
    double sum = a + b;
    return sum > 0.0 ? 2.0 * a * b / sum : 0.0;
}

// Rust's f64::max: the non-NaN operand when one is NaN.
__host__ __device__ __forceinline__ double rmax(double a, double b) { return std::fmax(a, b); }

// Wave-level butterfly sum: a fixed pairing, so the result depends only on the inputs.
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Block sum in a fixed order (waves by index); every thread passes its value, thread 0 gets the
// total.  `scratch` needs one double per wave.
template <int THREADS>
__device__ __forceinline__ double block_sum(double x, double *scratch) {
    constexpr int WAVES = THREADS / 64;
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    x = wave_sum(x);
    if ((t & 63) == 0) scratch[t >> 6] = x;
    __syncthreads();
    double total = 0.0;
    if (t == 0) {
#pragma unroll
        for (int w = 0; w < WAVES; w++) total += scratch[w];
    }
    return total;
}

// Sums n partials in a fixed order with one block of 1024 threads: thread t folds partials
// t, t+1024, ... in index order, then the block tree.  Deterministic for a given n.
__device__ __forceinline__ double fixed_order_sum_1024(const double *partials, int64_t n,
                                                       double *scratch) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) acc += partials[i];
    return block_sum<1024>(acc, scratch);
}

// Ledger::with over the given entries then audit_with against the previous total
// (conserved.rs:78-89, 323-363), on the device.  Called by thread 0 of a finalize kernel.
__device__ __forceinline__ void audit_step(AuditState *st, double total, double scale_entries,
                                           int64_t step, double *totals) {
    double b = st->prev_total, a = total;
    double scale = rmax(rmax(rmax(std::fabs(b), std::fabs(a)), st->prev_scale), scale_entries);
    if (!(scale < 1e-300) && std::fabs(a - b) / scale > st->tolerance) {
        st->halted = 1;
        st->breach_step = step;
        st->before = b, st->after = a, st->scale = scale;
    }
    st->prev_total = a;
    st->prev_scale = scale_entries;
    if (totals) totals[step] = a;
}

}  // namespace pm
