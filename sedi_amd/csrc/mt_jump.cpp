// mt_jump.cpp -- jump-ahead for mt19937 (host side).
//
// The probe stream std::mt19937(seed) is linear over GF(2): its 19937-bit
// state advances by a fixed linear map T whose characteristic polynomial phi
// has degree 19937 and is primitive.  So T^J = q(T) with q = x^J mod phi, and
// the state J outputs ahead is sum_{i: q_i = 1} T^i s (Haramoto, Matsumoto,
// Nishimura, L'Ecuyer, Panneton, "Efficient jump ahead for F2-linear random
// number generators", INFORMS J. Computing 20(3), 2008).
//
//   * phi comes from the Berlekamp-Massey algorithm on 2*19937 output bits
//     (any nonzero linear functional of the state has minimal polynomial phi);
//   * x^J mod phi by square-and-multiply on 19937-bit polynomials;
//   * T^i s in the order oldest-word-first is the window (w_i .. w_{i+623}) of
//     the untempered word sequence w started from s, so the jumped state is
//     the correlation acc[j] = XOR_{i: q_i = 1} w_{i+j}, j = 0..623.
//
// A state here is the 624 words oldest first, with the next output being the
// tempered value of the word generated after them (MtState::idx = 624).
#include "mt_jump.h"

#include <cstring>
#include <map>
#include <mutex>

namespace sedi {

namespace {
constexpr int N = 624, M = 397, DEG = 19937;
constexpr int W = (DEG + 64) / 64 + 1;  // words per reduced polynomial (313)

inline uint32_t next_word(uint32_t a, uint32_t b, uint32_t far) {
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

inline uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// Untempered words w_0 .. w_{count-1} starting from state s (w_0..w_623 = s).
void word_sequence(const uint32_t* s, std::vector<uint32_t>& w, size_t count) {
    w.resize(count);
    for (int i = 0; i < N && i < (int)count; i++) w[i] = s[i];
    for (size_t n = N; n < count; n++) w[n] = next_word(w[n - N], w[n - N + 1], w[n - N + M]);
}

using Poly = std::vector<uint64_t>;  // bit i = coefficient of x^i

inline bool getbit(const Poly& p, size_t i) { return (p[i >> 6] >> (i & 63)) & 1u; }

// phi, from Berlekamp-Massey on the low bit of successive outputs of a seeded
// generator.
Poly compute_charpoly() {
    const int len = 2 * DEG;
    uint32_t s[N];
    mt_seed_state(5489u, s);
    std::vector<uint32_t> w;
    word_sequence(s, w, (size_t)N + len);
    std::vector<uint8_t> bits(len);
    for (int n = 0; n < len; n++) bits[n] = temper(w[N + n]) & 1u;

    const int NW = (len + 64) / 64 + 1;
    Poly C(NW, 0), B(NW, 0), T(NW, 0), R(NW, 0);
    C[0] = B[0] = 1;
    int L = 0, m = 1;
    for (int n = 0; n < len; n++) {
        // R bit j = s_{n-j}
        for (int k = NW - 1; k > 0; k--) R[k] = (R[k] << 1) | (R[k - 1] >> 63);
        R[0] = (R[0] << 1) | bits[n];
        uint64_t acc = 0;
        const int lw = L / 64 + 1;
        for (int k = 0; k < lw && k < NW; k++) acc ^= C[k] & R[k];
        // only coefficients 0..L of C are live; they are zero above L
        const int d = __builtin_parityll(acc);
        if (!d) { m++; continue; }
        T = C;
        // C ^= B << m
        const int ws = m / 64, bs = m % 64;
        for (int k = NW - 1; k >= ws; k--) {
            uint64_t v = B[k - ws] << bs;
            if (bs && k - ws - 1 >= 0) v |= B[k - ws - 1] >> (64 - bs);
            C[k] ^= v;
        }
        if (2 * L <= n) {
            L = n + 1 - L;
            B = T;
            m = 1;
        } else {
            m++;
        }
    }
    // phi(x) = x^L C(1/x)
    Poly phi(W, 0);
    if (L != DEG) return Poly();  // cannot happen for mt19937
    for (int i = 0; i <= L; i++)
        if (getbit(C, i)) phi[(L - i) >> 6] |= 1ull << ((L - i) & 63);
    return phi;
}

const Poly& charpoly() {
    static std::once_flag once;
    static Poly phi;
    std::call_once(once, [] { phi = compute_charpoly(); });
    return phi;
}

// r = r mod phi for deg(r) < 2*DEG
void reduce(Poly& r, const Poly& phi) {
    for (long d = (long)r.size() * 64 - 1; d >= DEG; d--) {
        if (!getbit(r, d)) continue;
        const long sh = d - DEG;
        const long ws = sh / 64, bs = sh % 64;
        for (int k = 0; k < W; k++) {
            const uint64_t v = phi[k];
            if (!v) continue;
            r[k + ws] ^= v << bs;
            if (bs && k + ws + 1 < (long)r.size()) r[k + ws + 1] ^= v >> (64 - bs);
        }
    }
    r.resize(W);
}

void square_mod(Poly& p, const Poly& phi) {
    Poly sq(2 * W + 2, 0);
    for (int k = 0; k < W; k++) {
        uint64_t v = p[k];
        uint64_t lo = 0, hi = 0;
        for (int b = 0; b < 32; b++) {
            lo |= ((v >> b) & 1ull) << (2 * b);
            hi |= ((v >> (b + 32)) & 1ull) << (2 * b);
        }
        sq[2 * k] = lo;
        sq[2 * k + 1] = hi;
    }
    reduce(sq, phi);
    p.swap(sq);
}

// carry-less 64x64 -> 128 multiply (PCLMULQDQ on the host)
__attribute__((target("pclmul,sse2"))) inline void clmul64(uint64_t a, uint64_t b, uint64_t* lo,
                                                           uint64_t* hi) {
    typedef long long v2di __attribute__((vector_size(16)));
    v2di va = {(long long)a, 0}, vb = {(long long)b, 0};
    v2di r = __builtin_ia32_pclmulqdq128(va, vb, 0x00);
    *lo = (uint64_t)r[0];
    *hi = (uint64_t)r[1];
}

// r = a * b mod phi
Poly mul_mod(const Poly& a, const Poly& b, const Poly& phi) {
    Poly prod(2 * W + 2, 0);
    for (int i = 0; i < W; i++) {
        if (!a[i]) continue;
        for (int j = 0; j < W; j++) {
            if (!b[j]) continue;
            uint64_t lo, hi;
            clmul64(a[i], b[j], &lo, &hi);
            prod[i + j] ^= lo;
            prod[i + j + 1] ^= hi;
        }
    }
    reduce(prod, phi);
    return prod;
}

void mulx_mod(Poly& p, const Poly& phi) {
    uint64_t carry = 0;
    for (int k = 0; k < W; k++) {
        const uint64_t nc = p[k] >> 63;
        p[k] = (p[k] << 1) | carry;
        carry = nc;
    }
    if (getbit(p, DEG))
        for (int k = 0; k < W; k++) p[k] ^= phi[k];
}
}  // namespace

void mt_seed_state(uint32_t seed, uint32_t* s) {
    s[0] = seed;
    for (int i = 1; i < N; i++) s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
}

const std::vector<uint64_t>& mt_jump_poly(unsigned long long J) {
    static std::mutex mu;
    static std::map<unsigned long long, Poly> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(J);
    if (it != cache.end()) return it->second;
    const Poly& phi = charpoly();
    Poly r(W, 0);
    r[0] = 1;
    int top = 63;
    while (top >= 0 && !((J >> top) & 1ull)) top--;
    for (int b = top; b >= 0; b--) {
        square_mod(r, phi);
        if ((J >> b) & 1ull) mulx_mod(r, phi);
    }
    return cache.emplace(J, std::move(r)).first->second;
}

const std::vector<uint64_t>& mt_probe_poly(long long ncells, int k) {
    static std::mutex mu;
    static std::map<std::pair<long long, int>, Poly> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find({ncells, k});
        if (it != cache.end()) return it->second;
    }
    Poly r;
    if (k == 0) {
        r.assign(W, 0);
        r[0] = 1;
    } else {
        const Poly& step = mt_jump_poly(2ull * (unsigned long long)ncells);
        const Poly& prev = mt_probe_poly(ncells, k - 1);
        r = mul_mod(prev, step, charpoly());
    }
    std::lock_guard<std::mutex> lk(mu);
    return cache.emplace(std::make_pair(ncells, k), std::move(r)).first->second;
}

int mt_poly_words() { return W; }

void mt_poly_powers(unsigned long long J, int count, std::vector<uint64_t>& out) {
    static std::mutex mu;
    static std::map<std::pair<unsigned long long, int>, std::vector<uint64_t>> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find({J, count});
        if (it != cache.end()) {
            out = it->second;
            return;
        }
    }
    std::vector<uint64_t> all((size_t)W * count, 0);
    const Poly& step = mt_jump_poly(J);
    Poly cur(W, 0);
    cur[0] = 1;
    for (int j = 0; j < count; j++) {
        std::memcpy(all.data() + (size_t)W * j, cur.data(), sizeof(uint64_t) * W);
        if (j + 1 < count) cur = mul_mod(cur, step, charpoly());
    }
    std::lock_guard<std::mutex> lk(mu);
    out = cache.emplace(std::make_pair(J, count), std::move(all)).first->second;
}

void mt_word_sequence(const uint32_t* s, uint32_t* w, int count) {
    std::vector<uint32_t> v;
    word_sequence(s, v, (size_t)count);
    std::memcpy(w, v.data(), sizeof(uint32_t) * count);
}

void mt_jump_state(const uint32_t* s, unsigned long long J, uint32_t* out) {
    if (J == 0) {
        std::memcpy(out, s, sizeof(uint32_t) * N);
        return;
    }
    const Poly& q = mt_jump_poly(J);
    std::vector<uint32_t> w;
    word_sequence(s, w, (size_t)DEG + N);
    uint32_t acc[N];
    std::memset(acc, 0, sizeof(acc));
    for (int i = 0; i < DEG; i++) {
        if (!getbit(q, i)) continue;
        const uint32_t* src = w.data() + i;
        for (int j = 0; j < N; j++) acc[j] ^= src[j];
    }
    std::memcpy(out, acc, sizeof(acc));
}

bool mt_charpoly_ok() { return !charpoly().empty(); }

}  // namespace sedi
