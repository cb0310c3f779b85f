// rng.hip -- Rademacher probes from a device-resident mt19937 stream.
//
// The reference fills each probe with std::mt19937 and
// std::uniform_real_distribution<double>(0,1), +1 iff the draw is > 0.5, cells
// in i3,i2,i1 order (DistrRngStd.hpp:27-41).  Here the stream is
// std::mt19937(seed) per run: cell c of sample k takes 32-bit outputs 2Nk+2c
// (x0) and 2Nk+2c+1 (x1) and the draw is libstdc++'s generate_canonical:
// (x0 + x1 * 2^32) / 2^64 in double arithmetic.
//
// mt19937 (Matsumoto & Nishimura 1998) regenerates its 624-word state in
// place; word i depends on words i, i+1 and i+397 (mod 624) of the state as
// it stands, so a block regeneration splits into three dependent phases of
// 227, 227 and 170 words.  One workgroup keeps the state in LDS, runs the
// three phases with a barrier each (double-buffered state, so a phase never
// reads a word it overwrites), tempers the new words as it produces them and
// turns each pair of outputs into a probe value.  The stream is sequential by
// nature; the estimator runs it on its own HIP stream, one probe ahead of the
// solver, on a single CU.
#include "rng.h"
#include "mt_jump.h"
#include "profiler.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace sedi {

namespace {
constexpr int MT_N = 624, MT_M = 397;
constexpr uint32_t MT_A = 0x9908b0dfu, UPPER = 0x80000000u, LOWER = 0x7fffffffu;

__device__ __forceinline__ uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ uint32_t twist(uint32_t cur, uint32_t next, uint32_t far) {
    const uint32_t y = (cur & UPPER) | (next & LOWER);
    return far ^ (y >> 1) ^ ((y & 1u) ? MT_A : 0u);
}

// libstdc++ generate_canonical<double,53> on two outputs, then `> 0.5`.
__device__ __forceinline__ double rademacher(uint32_t x0, uint32_t x1) {
    double sum = (double)x0 * 1.0;
    sum = sum + (double)x1 * 4294967296.0;
    double ret = sum / 18446744073709551616.0;
    if (ret >= 1.0) ret = 0.99999999999999989;  // nextafter(1, 0)
    return ret > 0.5 ? 1.0 : -1.0;
}
}  // namespace

__global__ void mt_seed_kernel(MtState* s, uint32_t seed) {
    if (threadIdx.x == 0) {
        uint32_t v = seed;
        s->mt[0] = v;
        for (int i = 1; i < MT_N; i++) {
            v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
            s->mt[i] = v;
        }
        s->idx = MT_N;
        s->carry_valid = 0;
        s->carry = 0;
    }
}

// Produces `ncells` probe values from the stream (or only advances it by
// `skip` outputs when out == nullptr).  One workgroup of 256 lanes.
__global__ __launch_bounds__(256) void mt_fill_kernel(MtState* s, double* out, long long ncells,
                                                      unsigned long long skip) {
    __shared__ uint32_t S[2][MT_N];   // state, double-buffered by block
    __shared__ uint32_t O[2][MT_N];   // tempered outputs of the block
    const int t = threadIdx.x;
    for (int i = t; i < MT_N; i += 256) S[0][i] = s->mt[i];
    int idx = s->idx;
    int carry_valid = s->carry_valid;
    uint32_t carry = s->carry;
    __syncthreads();
    int cur = 0;
    if (idx < MT_N)
        for (int i = t; i < MT_N; i += 256) O[0][i] = temper(S[0][i]);
    __syncthreads();

    long long done = 0;  // cells written
    unsigned long long skipped = 0;
    const bool skipping = (out == nullptr);
    while (skipping ? (skipped < skip) : (done < ncells)) {
        if (idx >= MT_N) {
            // regenerate: three phases, reading S[cur], writing S[nxt] and O[nxt]
            const int nxt = cur ^ 1;
            if (t < 227) {
                const uint32_t v = twist(S[cur][t], S[cur][t + 1], S[cur][t + MT_M]);
                S[nxt][t] = v;
                O[nxt][t] = temper(v);
            }
            __syncthreads();
            if (t < 227) {
                const int i = t + 227;
                const uint32_t v = twist(S[cur][i], S[cur][i + 1], S[nxt][i - 227]);
                S[nxt][i] = v;
                O[nxt][i] = temper(v);
            }
            __syncthreads();
            if (t < 170) {
                const int i = t + 454;
                const uint32_t nx = (i + 1 < MT_N) ? S[cur][i + 1] : S[nxt][0];
                const uint32_t v = twist(S[cur][i], nx, S[nxt][i - 227]);
                S[nxt][i] = v;
                O[nxt][i] = temper(v);
            }
            __syncthreads();
            cur = nxt;
            idx = 0;
        }
        if (skipping) {
            const unsigned long long avail = (unsigned long long)(MT_N - idx);
            const unsigned long long take = (skip - skipped) < avail ? (skip - skipped) : avail;
            idx += (int)take;
            skipped += take;
            continue;
        }
        const uint32_t* o = O[cur];
        if (carry_valid) {
            if (t == 0) out[done] = rademacher(carry, o[idx]);
            idx += 1;
            done += 1;
            carry_valid = 0;
            continue;
        }
        long long pairs = (MT_N - idx) / 2;
        if (pairs > ncells - done) pairs = ncells - done;
        for (int p = t; p < pairs; p += 256) out[done + p] = rademacher(o[idx + 2 * p], o[idx + 2 * p + 1]);
        idx += 2 * (int)pairs;
        done += pairs;
        if (done < ncells && idx == MT_N - 1) {
            carry = o[idx];
            carry_valid = 1;
            idx = MT_N;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int i = t; i < MT_N; i += 256) s->mt[i] = S[cur][i];
    if (t == 0) {
        s->idx = idx;
        s->carry_valid = carry_valid;
        s->carry = carry;
    }
}

// ---------------------------------------------------------------------------
// Batched probes, in parallel.  Probe k of the stream starts 2Nk outputs
// after the stream's start; with q = x^J mod phi the state J outputs after a
// state s has word j = XOR of w_{i+j} over the set bits i of q, where w is the
// untempered word sequence from s (Haramoto et al. 2008, see mt_jump.cpp).
//   1. mt_probe_start_kernel, one workgroup per probe: jumps the stream's
//      start (whose word sequence is uploaded once per seed) by 2Nk and
//      writes the probe's own word sequence (MT_WINDOW words);
//   2. mt_probe_seg_kernel, one workgroup per (segment, probe): jumps that
//      sequence by 2*S*j to the start of segment j (S cells; the polynomials
//      are seed-independent and cached per grid size) and generates the
//      segment like mt_fill_kernel, into lane k%4 of RHS block k/4 of the
//      batch layout [K/4][N][4].
// Segments start at even outputs, so a cell's two outputs never straddle.
// ---------------------------------------------------------------------------

// The state `q` jumps to, from the word sequence w: S[j] = XOR_{q_i = 1} w[i+j].
__device__ __forceinline__ void jump_state(const uint32_t* __restrict__ w, const uint64_t* __restrict__ q,
                                           int pw, uint32_t* S) {
    const int t = threadIdx.x;
    uint32_t a0 = 0, a1 = 0, a2 = 0;
    for (int wi = 0; wi < pw; wi++) {
        uint64_t bits = q[wi];
        while (bits) {
            const int b = __builtin_ctzll(bits);
            bits &= bits - 1;
            const uint32_t* src = w + wi * 64 + b;
            a0 ^= src[t];
            a1 ^= src[t + 256];
            if (t < MT_N - 512) a2 ^= src[t + 512];
        }
    }
    S[t] = a0;
    S[t + 256] = a1;
    if (t < MT_N - 512) S[t + 512] = a2;
}

// One regeneration of the 624-word state S[cur] into S[nxt] (and, if O is
// given, the tempered outputs into O[nxt]); three dependent phases.
__device__ __forceinline__ void regenerate(uint32_t (*S)[MT_N], uint32_t (*O)[MT_N], int cur) {
    const int t = threadIdx.x;
    const int nxt = cur ^ 1;
    if (t < 227) {
        const uint32_t v = twist(S[cur][t], S[cur][t + 1], S[cur][t + MT_M]);
        S[nxt][t] = v;
        if (O) O[nxt][t] = temper(v);
    }
    __syncthreads();
    if (t < 227) {
        const int i = t + 227;
        const uint32_t v = twist(S[cur][i], S[cur][i + 1], S[nxt][i - 227]);
        S[nxt][i] = v;
        if (O) O[nxt][i] = temper(v);
    }
    __syncthreads();
    if (t < 170) {
        const int i = t + 454;
        const uint32_t nx = (i + 1 < MT_N) ? S[cur][i + 1] : S[nxt][0];
        const uint32_t v = twist(S[cur][i], nx, S[nxt][i - 227]);
        S[nxt][i] = v;
        if (O) O[nxt][i] = temper(v);
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void mt_probe_start_kernel(const uint32_t* __restrict__ wseq,
                                                             const uint64_t* __restrict__ polys, int pw,
                                                             uint32_t* __restrict__ wk) {
    __shared__ uint32_t S[2][MT_N];
    const int k = blockIdx.x;
    uint32_t* out = wk + (long long)k * MT_WINDOW;
    jump_state(wseq, polys + (long long)k * pw, pw, S[0]);
    __syncthreads();
    for (int i = threadIdx.x; i < MT_N; i += 256) out[i] = S[0][i];
    int cur = 0;
    for (int produced = MT_N; produced < MT_WINDOW; produced += MT_N) {
        regenerate(S, nullptr, cur);
        cur ^= 1;
        for (int i = threadIdx.x; i < MT_N && produced + i < MT_WINDOW; i += 256) out[produced + i] = S[cur][i];
    }
}

template <class T>
__global__ __launch_bounds__(256) void mt_probe_seg_kernel(const uint32_t* __restrict__ wk,
                                                           const uint64_t* __restrict__ segpolys, int pw,
                                                           long long ncells, long long seg_cells, int nprobe,
                                                           T* __restrict__ out) {
    __shared__ uint32_t S[2][MT_N];
    __shared__ uint32_t O[2][MT_N];
    const int t = threadIdx.x;
    const int j = blockIdx.x, k = blockIdx.y;
    const long long cell0 = (long long)j * seg_cells;
    if (cell0 >= ncells) return;
    const long long cells = min(seg_cells, ncells - cell0);
    T* dst = out + (long long)(k >> 2) * ncells * 4 + (k & 3) + cell0 * 4;
    if (k >= nprobe) {  // padding: a zero right-hand side
        for (long long i = t; i < cells; i += 256) dst[i * 4] = (T)0;
        return;
    }
    jump_state(wk + (long long)k * MT_WINDOW, segpolys + (long long)j * pw, pw, S[0]);
    __syncthreads();
    int cur = 0;
    long long done = 0;
    while (done < cells) {
        regenerate(S, O, cur);
        cur ^= 1;
        // 312 cells per block of outputs
        long long pairs = MT_N / 2;
        if (pairs > cells - done) pairs = cells - done;
        for (int p = t; p < pairs; p += 256)
            dst[(done + p) * 4] = (T)rademacher(O[cur][2 * p], O[cur][2 * p + 1]);
        done += pairs;
    }
}

namespace {
struct SegPlan {
    long long seg_cells = 0;
    int nseg = 0;
    DevBuf polys;  // nseg x pw words: x^(2*seg_cells*j) mod phi
};

// Segment plan of a grid size: at least 65536 cells and at most 256
// segments per probe; built once per process.
const SegPlan* seg_plan(long long ncells, int* rc) {
    static std::mutex mu;
    static std::map<long long, std::unique_ptr<SegPlan>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(ncells);
    if (it != cache.end()) return it->second.get();
    std::unique_ptr<SegPlan> p(new SegPlan());
    p->seg_cells = std::max<long long>(65536, (ncells + 255) / 256);
    p->nseg = (int)((ncells + p->seg_cells - 1) / p->seg_cells);
    const int pw = mt_poly_words();
    std::vector<uint64_t> host;
    mt_poly_powers(2ull * (unsigned long long)p->seg_cells, p->nseg, host);
    if (p->polys.reserve(sizeof(uint64_t) * host.size()) != hipSuccess ||
        hipMemcpy(p->polys.p, host.data(), sizeof(uint64_t) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("cannot upload the probe segment polynomials");
        *rc = SEDI_ERR_HIP;
        return nullptr;
    }
    (void)pw;
    return cache.emplace(ncells, std::move(p)).first->second.get();
}
}  // namespace

int ProbeSource::init(uint32_t seed, long long ncells, int k0, int count, hipStream_t stream) {
    SEDI_REQUIRE(ncells > 0 && k0 >= 0 && count >= 0, SEDI_ERR_ARG, "bad probe source (k0 %d, count %d)", k0,
                 count);
    SEDI_REQUIRE(mt_charpoly_ok(), SEDI_ERR_STATE, "mt19937 characteristic polynomial unavailable");
    ncells_ = ncells;
    count_ = count;
    if (count == 0) return SEDI_OK;
    int rc = SEDI_OK;
    if (!seg_plan(ncells, &rc)) return rc;
    const int pw = mt_poly_words();
    const size_t wbytes = sizeof(uint32_t) * MT_WINDOW;
    const size_t pbytes = sizeof(uint64_t) * (size_t)pw * count;
    std::vector<uint8_t> host(wbytes + pbytes);
    uint32_t s0[624];
    mt_seed_state(seed, s0);
    mt_word_sequence(s0, reinterpret_cast<uint32_t*>(host.data()), MT_WINDOW);
    for (int k = 0; k < count; k++) {
        const std::vector<uint64_t>& q = mt_probe_poly(ncells, k0 + k);
        std::memcpy(host.data() + wbytes + sizeof(uint64_t) * (size_t)pw * k, q.data(), sizeof(uint64_t) * pw);
    }
    SEDI_HIP(up_.reserve(host.size()));
    SEDI_HIP(wk_.reserve(sizeof(uint32_t) * (size_t)MT_WINDOW * count));
    // the host staging vector must outlive the copy: a synchronous one
    SEDI_HIP(hipMemcpy(up_.p, host.data(), host.size(), hipMemcpyHostToDevice));
    const uint32_t* wseq = up_.as<uint32_t>();
    const uint64_t* polys = reinterpret_cast<const uint64_t*>(static_cast<char*>(up_.p) + wbytes);
    SEDI_HIP(launch("mt_probe_start_kernel", mt_probe_start_kernel, dim3(count), dim3(256), stream, wseq, polys,
                    pw, wk_.as<uint32_t>()));
    return SEDI_OK;
}

int ProbeSource::emit(int first, int nprobe, int K, bool fp32, void* out, hipStream_t stream) const {
    SEDI_REQUIRE(K > 0 && K % 4 == 0 && first >= 0 && nprobe >= 0 && nprobe <= K && first + nprobe <= count_,
                 SEDI_ERR_ARG, "bad probe emit (first %d, count %d of %d, K %d)", first, nprobe, count_, K);
    int rc = SEDI_OK;
    const SegPlan* plan = seg_plan(ncells_, &rc);
    if (!plan) return rc;
    const uint32_t* wk = wk_.as<uint32_t>() + (size_t)MT_WINDOW * first;
    const dim3 grid(plan->nseg, K);
    if (fp32)
        SEDI_HIP(launch("mt_probe_seg_kernel", mt_probe_seg_kernel<float>, grid, dim3(256), stream, wk,
                        plan->polys.as<uint64_t>(), mt_poly_words(), ncells_, plan->seg_cells, nprobe,
                        static_cast<float*>(out)));
    else
        SEDI_HIP(launch("mt_probe_seg_kernel", mt_probe_seg_kernel<double>, grid, dim3(256), stream, wk,
                        plan->polys.as<uint64_t>(), mt_poly_words(), ncells_, plan->seg_cells, nprobe,
                        static_cast<double*>(out)));
    return SEDI_OK;
}

int fill_probe_batch(uint32_t seed, long long ncells, int k0, int K, int kact, bool fp32, void* out,
                     hipStream_t stream, DevBuf& scratch) {
    SEDI_REQUIRE(K > 0 && K % 4 == 0 && kact >= 0 && kact <= K && k0 >= 0, SEDI_ERR_ARG,
                 "bad probe batch (k0 %d, K %d, active %d)", k0, K, kact);
    (void)scratch;
    ProbeSource src;
    int rc = src.init(seed, ncells, k0, kact, stream);
    if (!rc) rc = src.emit(0, kact, K, fp32, out, stream);
    if (rc) return rc;
    // the source's buffers are released on return
    SEDI_HIP(hipStreamSynchronize(stream));
    return SEDI_OK;
}

__global__ void unbatch_rows_kernel(const double* __restrict__ batch, long long ncells, int count,
                                    double* __restrict__ rows) {
    const long long total = ncells * count;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long k = e / ncells, i = e - k * ncells;
        rows[e] = batch[(k >> 2) * ncells * 4 + i * 4 + (k & 3)];
    }
}

int unbatch_rows(const double* batch, long long ncells, int count, double* rows, hipStream_t stream) {
    long long g = (ncells * count + 255) / 256;
    if (g > 8192) g = 8192;
    SEDI_HIP(launch("unbatch_rows_kernel", unbatch_rows_kernel, dim3((unsigned)g), dim3(256), stream, batch,
                    ncells, count, rows));
    return SEDI_OK;
}

int MtStream::init() {
    if (state_.p) return SEDI_OK;
    SEDI_HIP(state_.reserve(sizeof(MtState)));
    return SEDI_OK;
}

int MtStream::seed(uint32_t seed, hipStream_t stream) {
    int rc = init();
    if (rc) return rc;
    SEDI_HIP(launch("mt_seed_kernel", mt_seed_kernel, dim3(1), dim3(64), stream, state_.as<MtState>(), seed));
    return SEDI_OK;
}

int MtStream::seed_at(uint32_t seed, unsigned long long first_draw, hipStream_t stream) {
    int rc = init();
    if (rc) return rc;
    MtState h;
    uint32_t s0[624];
    mt_seed_state(seed, s0);
    mt_jump_state(s0, first_draw, h.mt);
    h.idx = 624;
    h.carry_valid = 0;
    h.carry = 0;
    SEDI_HIP(hipMemcpyAsync(state_.p, &h, sizeof(h), hipMemcpyHostToDevice, stream));
    SEDI_HIP(hipStreamSynchronize(stream));
    return SEDI_OK;
}

int MtStream::skip(unsigned long long draws, hipStream_t stream) {
    if (!draws) return SEDI_OK;
    SEDI_HIP(launch("mt_fill_kernel", mt_fill_kernel, dim3(1), dim3(256), stream, state_.as<MtState>(),
                    (double*)nullptr, 0LL, draws));
    return SEDI_OK;
}

int MtStream::fill(double* v, long long ncells, hipStream_t stream) {
    if (ncells <= 0) return SEDI_OK;
    SEDI_HIP(launch("mt_fill_kernel", mt_fill_kernel, dim3(1), dim3(256), stream, state_.as<MtState>(), v,
                    ncells, 0ULL));
    return SEDI_OK;
}

}  // namespace sedi
