// nrq_core.hip -- error handling, the library stream, the lattice module and
// the host ranlxd generator.
#include "nrq_internal.h"
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace nrq {

static int g_last_code = 0;
static char g_last_msg[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_msg, sizeof(g_last_msg), fmt, ap);
    va_end(ap);
    g_last_code = code;
    return code;
}

int check_hip(hipError_t e, const char *what)
{
    return fail(NRQ_ERR_HIP, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
}

lattice_state &lat()
{
    static lattice_state s;
    return s;
}

static hipStream_t g_stream = nullptr;
static bool g_device_ok = false;

int ensure_device()
{
    if (g_device_ok) return NRQ_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1)
        return fail(NRQ_ERR_HIP, "no HIP device available (hipGetDeviceCount: %s, count %d)",
                    hipGetErrorString(e), n);
    NRQ_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device_ok = true;
    return NRQ_OK;
}

hipStream_t stream() { return g_stream; }

int scratch_buf::reserve(size_t n)
{
    if (n <= bytes) return NRQ_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(NRQ_ERR_ALLOC, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return NRQ_OK;
}

scratch_buf::~scratch_buf()
{
    // Process teardown: the runtime may already be gone; leak rather than
    // call into it.
}

// ---------------------------------------------------------------------------
// ranlxd (Luscher's RANLUX, double precision, 48-bit base; the algorithm of
// ranlxd_generator.f90:77-349).  Each update() applies pr steps of the
// subtract-with-borrow recurrence x[ir] <- x[jr] - x[ir] - carry on 12
// numbers that are exact multiples of 2^-48 in [0, 1); 12 numbers are handed
// out per update.  The numbers are held as their 48-bit integer numerators:
// every step of the double-precision recurrence is exact, so the integer form
// gives bit-identical output, and the borrow comes from the sign bit instead
// of a data-dependent branch.
// ---------------------------------------------------------------------------
namespace {
struct Ranlux48 {
    int64_t x[12];   // numerators of the state, x / 2^48 in [0, 1)
    int64_t carry = 0; // 0 or 1 (units of 2^-48)
    int ir = 0, jr = 0, ir_old = 0;
    int64_t pr = 0;
    bool ready = false;

    int seed(int level, int64_t s)
    {
        if (level != 1 && level != 2) return -1;
        if (s <= 0 || s > 2147483647LL) return -2;
        pr = level == 1 ? 202 : 397;
        int bit[31];
        for (int k = 0; k < 31; ++k) bit[k] = (int)((s >> k) & 1);
        int a = 0, b = 18;
        for (int k = 0; k < 12; ++k) {
            int64_t acc = 0;
            for (int l = 0; l < 48; ++l) {
                acc = 2 * acc + (1 - bit[a]);
                bit[a] ^= bit[b];
                a = a == 30 ? 0 : a + 1;
                b = b == 30 ? 0 : b + 1;
            }
            x[k] = acc;
        }
        carry = 0;
        ir = 11;
        jr = 7;
        ir_old = 0;
        ready = true;
        return 0;
    }

    // one step of the recurrence at (i, j); returns the new carry
    static inline int64_t step(int64_t *x, int i, int j, int64_t c)
    {
        constexpr int64_t ONE = int64_t(1) << 48;
        const int64_t y = x[j] - x[i] - c;
        const int64_t borrow = y >> 63; // -1 if y < 0, else 0
        x[i] = y + (borrow & ONE);
        return borrow & 1;
    }

    void refresh()
    {
        int64_t c = carry;
        int i = ir, j = jr;
        int64_t k = 0;
        // single steps up to i == 0, then whole turns of the 12 positions
        // with constant indices (j - i = 8 mod 12 throughout), then the rest
        for (; k < pr && i != 0; ++k) {
            c = step(x, i, j, c);
            i = i == 11 ? 0 : i + 1;
            j = j == 11 ? 0 : j + 1;
        }
        if (i == 0 && j == 8) {
            for (; k + 12 <= pr; k += 12) {
#pragma unroll
                for (int q = 0; q < 12; ++q) c = step(x, q, (q + 8) % 12, c);
            }
        }
        for (; k < pr; ++k) {
            c = step(x, i, j, c);
            i = i == 11 ? 0 : i + 1;
            j = j == 11 ? 0 : j + 1;
        }
        carry = c;
        ir = i;
        jr = j;
        ir_old = ir;
    }

    double next()
    {
        constexpr double UNIT = 1.0 / 281474976710656.0; // 2^-48, exact scaling
        ir = ir == 11 ? 0 : ir + 1;
        if (ir == ir_old) refresh();
        return (double)x[ir] * UNIT;
    }
};
Ranlux48 g_rng;
} // namespace

int rng_init(int level, int64_t s)
{
    int r = g_rng.seed(level, s);
    if (r == -1) return fail(NRQ_ERR_ARG, "ranlxd: luxury level must be 1 or 2");
    if (r == -2) return fail(NRQ_ERR_ARG, "ranlxd: seed must be in [1, 2^31-1]");
    return NRQ_OK;
}

void rng_uniform(double *r, int64_t n)
{
    if (!g_rng.ready) g_rng.seed(1, 1); // ranlxd_generator.f90:328-330
    for (int64_t k = 0; k < n; ++k) r[k] = g_rng.next();
}

bool rng_ready() { return g_rng.ready; }

} // namespace nrq

using namespace nrq;

extern "C" int nrq_last_error(char *buf, size_t n)
{
    if (buf && n) {
        strncpy(buf, g_last_msg, n - 1);
        buf[n - 1] = 0;
    }
    return g_last_code;
}

extern "C" const char *nrq_version(void) { return "nrqcd_amd 0.1 (gfx950)"; }

extern "C" int nrq_lattice_init(const int64_t ext[3], const double spacing[4])
{
    if (lat().initialised) return fail(NRQ_ERR_STATE, "lattice: already initialised");
    if (!ext || !spacing) return fail(NRQ_ERR_ARG, "lattice: null argument");
    for (int i = 0; i < 3; ++i)
        if (ext[i] < 1 || ext[i] > (1 << 20)) return fail(NRQ_ERR_ARG, "lattice: extent %d = %lld out of range", i + 1, (long long)ext[i]);
    for (int i = 0; i < 4; ++i)
        if (!(spacing[i] > 0) || !std::isfinite(spacing[i]))
            return fail(NRQ_ERR_ARG, "lattice: spacing a%d = %g must be positive", i, spacing[i]);
    // kernels index sites and grids with 32-bit arithmetic
    if (ext[0] * ext[1] * ext[2] > (int64_t)INT32_MAX / 32)
        return fail(NRQ_ERR_ARG, "lattice: volume %lld exceeds %d sites", (long long)(ext[0] * ext[1] * ext[2]),
                    INT32_MAX / 32);
    geom &g = lat().g;
    g.n0 = (int32_t)ext[0];
    g.n1 = (int32_t)ext[1];
    g.n2 = (int32_t)ext[2];
    g.vol = ext[0] * ext[1] * ext[2];
    for (int i = 0; i < 4; ++i) g.a[i] = spacing[i];
    lat().initialised = true;
    return NRQ_OK;
}

extern "C" int nrq_lattice_finalize(void)
{
    if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised");
    gauge_forget_lattice();
    lat().initialised = false;
    return NRQ_OK;
}

#define NEED_LATTICE()                                                              \
    do {                                                                            \
        if (!lat().initialised) return fail(NRQ_ERR_STATE, "lattice: not initialised"); \
    } while (0)

static int check_idx(int64_t idx)
{
    if (idx < 1 || idx > lat().g.vol) return fail(NRQ_ERR_ARG, "lattice index %lld out of range", (long long)idx);
    return NRQ_OK;
}

extern "C" int nrq_lattice_size(int64_t *vol)
{
    NEED_LATTICE();
    if (!vol) return fail(NRQ_ERR_ARG, "null argument");
    *vol = lat().g.vol;
    return NRQ_OK;
}

extern "C" int nrq_lattice_index(const int64_t pos[3], int64_t *idx)
{
    NEED_LATTICE();
    const geom &g = lat().g;
    if (!pos || !idx) return fail(NRQ_ERR_ARG, "null argument");
    int64_t n[3] = {g.n0, g.n1, g.n2};
    for (int d = 0; d < 3; ++d)
        if (pos[d] < 1 || pos[d] > n[d]) return fail(NRQ_ERR_ARG, "position component %d out of range", d + 1);
    *idx = 1 + (pos[0] - 1) + n[0] * ((pos[1] - 1) + n[1] * (pos[2] - 1));
    return NRQ_OK;
}

static void position0(int64_t s, int64_t p[3])
{
    const geom &g = lat().g;
    p[0] = s % g.n0;
    p[1] = (s / g.n0) % g.n1;
    p[2] = s / ((int64_t)g.n0 * g.n1);
}

extern "C" int nrq_lattice_position(int64_t idx, int64_t pos[3])
{
    NEED_LATTICE();
    if (!pos) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = check_idx(idx)) return r;
    int64_t p[3];
    position0(idx - 1, p);
    for (int d = 0; d < 3; ++d) pos[d] = p[d] + 1;
    return NRQ_OK;
}

extern "C" int nrq_lattice_neib(int dir, int64_t idx, int64_t *out)
{
    NEED_LATTICE();
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    if (dir < -3 || dir > 3) return fail(NRQ_ERR_ARG, "direction %d out of range", dir);
    if (int r = check_idx(idx)) return r;
    if (dir == 0) {
        *out = idx;
        return NRQ_OK;
    }
    const geom &g = lat().g;
    int64_t n[3] = {g.n0, g.n1, g.n2};
    int64_t p[3];
    position0(idx - 1, p);
    int d = (dir > 0 ? dir : -dir) - 1;
    p[d] = (p[d] + (dir > 0 ? 1 : n[d] - 1)) % n[d];
    *out = 1 + p[0] + n[0] * (p[1] + n[1] * p[2]);
    return NRQ_OK;
}

extern "C" int nrq_lattice_negative_index(int64_t idx, int64_t *out)
{
    NEED_LATTICE();
    if (!out) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = check_idx(idx)) return r;
    const geom &g = lat().g;
    int64_t n[3] = {g.n0, g.n1, g.n2};
    int64_t p[3];
    position0(idx - 1, p);
    for (int d = 0; d < 3; ++d) p[d] = (n[d] - p[d]) % n[d];
    *out = 1 + p[0] + n[0] * (p[1] + n[1] * p[2]);
    return NRQ_OK;
}

extern "C" int nrq_lattice_momentum(int64_t idx, double pm[6])
{
    NEED_LATTICE();
    if (!pm) return fail(NRQ_ERR_ARG, "null argument");
    if (int r = check_idx(idx)) return r;
    const geom &g = lat().g;
    int64_t n[3] = {g.n0, g.n1, g.n2};
    int64_t p[3];
    position0(idx - 1, p);
    for (int d = 0; d < 3; ++d) {
        double arg = M_PI * (double)p[d] / (double)n[d];
        double mag = 2.0 * std::sin(arg) / g.a[d + 1];
        pm[2 * d] = mag * std::cos(arg);
        pm[2 * d + 1] = -mag * std::sin(arg);
    }
    return NRQ_OK;
}

extern "C" int nrq_lattice_norm2_momentum(int64_t idx, double *out)
{
    double pm[6];
    if (int r = nrq_lattice_momentum(idx, pm)) return r;
    double s = 0;
    for (int k = 0; k < 6; ++k) s += pm[k] * pm[k];
    *out = std::sqrt(s);
    return NRQ_OK;
}

extern "C" int nrq_random_init(int64_t seed) { return rng_init(2, seed); }

extern "C" int nrq_random_uniform(double *r, int64_t n)
{
    if (!r || n < 0) return fail(NRQ_ERR_ARG, "null argument or negative count");
    rng_uniform(r, n);
    return NRQ_OK;
}

// ---------------------------------------------------------------------------
// HIP events on the library stream, for timing from the host layer.
// ---------------------------------------------------------------------------
static hipEvent_t g_events[16];
static bool g_event_made[16];

extern "C" int nrq_event_record(int slot)
{
    if (slot < 0 || slot >= 16) return fail(NRQ_ERR_ARG, "event slot %d out of range", slot);
    if (int r = ensure_device()) return r;
    if (!g_event_made[slot]) {
        NRQ_HIP(hipEventCreate(&g_events[slot]));
        g_event_made[slot] = true;
    }
    NRQ_HIP(hipEventRecord(g_events[slot], stream()));
    return NRQ_OK;
}

extern "C" int nrq_event_elapsed(int start, int stop, double *ms)
{
    if (start < 0 || start >= 16 || stop < 0 || stop >= 16 || !ms) return fail(NRQ_ERR_ARG, "bad event slots");
    if (!g_event_made[start] || !g_event_made[stop]) return fail(NRQ_ERR_STATE, "event slot never recorded");
    NRQ_HIP(hipEventSynchronize(g_events[stop]));
    float t = 0;
    NRQ_HIP(hipEventElapsedTime(&t, g_events[start], g_events[stop]));
    *ms = t;
    return NRQ_OK;
}

extern "C" int nrq_synchronize(void)
{
    if (int r = ensure_device()) return r;
    NRQ_HIP(hipStreamSynchronize(stream()));
    return NRQ_OK;
}

// ---------------------------------------------------------------------------
// Streaming copy probe: the achievable HBM bandwidth of a plain read+write
// kernel on this device, and a known-size reference for hardware counters.
// ---------------------------------------------------------------------------
typedef double v2d __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_stream_copy(const v2d *__restrict__ src, v2d *__restrict__ dst, int64_t n)
{
    // each thread keeps four 16-byte loads in flight
    const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    v2d v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n) v[k] = __builtin_nontemporal_load(src + base + 256 * k);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n) __builtin_nontemporal_store(v[k], dst + base + 256 * k);
}

extern "C" int nrq_bandwidth_probe(int64_t bytes, int reps, double *gbs)
{
    if (bytes < 4096 || reps < 1 || !gbs) return fail(NRQ_ERR_ARG, "bandwidth probe: bad arguments");
    if (int r = ensure_device()) return r;
    const int64_t n = bytes / (int64_t)sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
        if (a) (void)hipFree(a);
        return fail(NRQ_ERR_ALLOC, "bandwidth probe: allocation of %lld bytes failed", (long long)bytes);
    }
    NRQ_HIP(hipMemsetAsync(a, 0, n * sizeof(double2), stream()));
    hipEvent_t e0, e1;
    NRQ_HIP(hipEventCreate(&e0));
    NRQ_HIP(hipEventCreate(&e1));
    const unsigned blocks = (unsigned)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, stream(), (const v2d *)a, (v2d *)b, n); // warm-up
    NRQ_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, stream(), (const v2d *)a, (v2d *)b, n);
    NRQ_HIP(hipEventRecord(e1, stream()));
    NRQ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NRQ_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    *gbs = 2.0 * (double)(n * sizeof(double2)) * reps / (ms * 1e-3) / 1e9;
    return NRQ_OK;
}
