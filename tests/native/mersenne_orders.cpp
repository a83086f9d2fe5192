// mersenne_orders.cpp -- every odd prime p < 2^B whose multiplicative order
// of 2 is below D, with that order, found without stepping any index.
//
// p | M_n = 2^n - 1  <=>  2^n = 1 (mod p)  <=>  ord_p(2) | n, so these pairs
// give the whole factor list of every Mersenne index n < D over p < 2^B, an
// independent check of a trial-factoring run of that size.  ord_p(2) = d
// implies d | p - 1 (and 2d | p - 1 for odd d, p being odd), so for each d
// the candidates are p = 1 + k s with s = d or 2d: each prime among them with
// 2^d = 1 (mod p) and 2^(d/q) != 1 for every prime q | d has order d.
//
//   mersenne_orders <B> <D> <out>   writes (u32 p, u32 d) pairs, sorted by p,
//                                   little-endian; B <= 31
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <omp.h>

namespace {

// odd n < 2^B: bit (n >> 1) of the bitmap is set when n is prime
std::vector<uint64_t> odd_prime_bitmap(uint32_t B)
{
    const uint64_t limit = 1ull << B;
    const uint64_t nodd = limit >> 1;
    std::vector<uint64_t> bits((nodd + 63) / 64, 0);
    uint32_t root = 1;
    while ((uint64_t)root * root < limit) ++root;
    std::vector<uint32_t> base;
    {
        std::vector<char> small(root + 1, 1);
        for (uint32_t i = 3; i <= root; i += 2)
            if (small[i]) {
                base.push_back(i);
                for (uint64_t j = (uint64_t)i * i; j <= root; j += 2 * i) small[j] = 0;
            }
    }
    const uint64_t SEG = 1ull << 18; // odd numbers per segment (a multiple of 64)
    const int64_t nseg = (int64_t)((nodd + SEG - 1) / SEG);
#pragma omp parallel
    {
        std::vector<char> comp(SEG);
#pragma omp for schedule(dynamic, 4)
        for (int64_t s = 0; s < nseg; ++s) {
            const uint64_t lo = (uint64_t)s * SEG; // odd index: n = 2 i + 1
            const uint64_t hi = std::min<uint64_t>(lo + SEG, nodd);
            std::fill(comp.begin(), comp.end(), 0);
            for (uint32_t q : base) {
                // first odd multiple of q that is >= max(q^2, 2 lo + 1)
                uint64_t start = std::max<uint64_t>((uint64_t)q * q, 2 * lo + 1);
                uint64_t m = (start + q - 1) / q * q;
                if (!(m & 1)) m += q;
                for (uint64_t i = (m - 1) / 2; i < hi; i += q) comp[i - lo] = 1;
            }
            if (lo == 0) comp[0] = 1; // 1 is not prime
            for (uint64_t i = lo; i < hi; i += 64) {
                uint64_t w = 0;
                for (uint64_t j = 0; j < 64 && i + j < hi; ++j)
                    if (!comp[i + j - lo]) w |= 1ull << j;
                bits[i / 64] = w;
            }
        }
    }
    return bits;
}

inline bool is_odd_prime(const std::vector<uint64_t> &bits, uint64_t n)
{
    const uint64_t i = n >> 1;
    return (bits[i >> 6] >> (i & 63)) & 1;
}

// Montgomery arithmetic modulo an odd p < 2^31 with R = 2^32 (t + m p < 2^64)
struct Mont32 {
    uint32_t p, pinv; // pinv = -p^-1 mod 2^32
    uint32_t one;     // R mod p
    explicit Mont32(uint32_t pp) : p(pp)
    {
        uint32_t x = pp; // Newton: x = p^-1 mod 2^32
        for (int i = 0; i < 5; ++i) x *= 2 - pp * x;
        pinv = 0u - x;
        one = (uint32_t)((1ull << 32) % pp);
    }
    uint32_t mul(uint32_t a, uint32_t b) const
    {
        const uint64_t t = (uint64_t)a * b;
        const uint32_t m = (uint32_t)t * pinv;
        const uint64_t u = (t + (uint64_t)m * p) >> 32; // < 2 p (t < p^2 < p 2^32)
        return (uint32_t)(u >= p ? u - p : u);
    }
    // 2^e in Montgomery form
    uint32_t pow2(uint32_t e) const
    {
        uint32_t r = one, x = one + one >= p ? one + one - p : one + one; // 2 R mod p
        while (e) {
            if (e & 1) r = mul(r, x);
            x = mul(x, x);
            e >>= 1;
        }
        return r;
    }
};

} // namespace

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: mersenne_orders <B> <D> <out>\n");
        return 2;
    }
    const uint32_t B = (uint32_t)std::strtoul(argv[1], nullptr, 10);
    const uint32_t D = (uint32_t)std::strtoul(argv[2], nullptr, 10);
    if (B < 3 || B > 31 || D < 2) {
        std::fprintf(stderr, "B must be in [3, 31] and D >= 2\n");
        return 2;
    }
    const uint64_t limit = 1ull << B;
    const std::vector<uint64_t> bits = odd_prime_bitmap(B);
    std::vector<std::pair<uint32_t, uint32_t>> out;
#pragma omp parallel
    {
        std::vector<std::pair<uint32_t, uint32_t>> mine;
#pragma omp for schedule(dynamic, 1)
        for (int64_t dd = 2; dd < (int64_t)D; ++dd) {
            const uint32_t d = (uint32_t)dd;
            std::vector<uint32_t> qs; // the prime factors of d
            for (uint32_t x = d, q = 2; x > 1; ++q) {
                if ((uint64_t)q * q > x) {
                    qs.push_back(x);
                    break;
                }
                if (x % q == 0) {
                    qs.push_back(q);
                    while (x % q == 0) x /= q;
                }
            }
            const uint64_t s = (d & 1) ? 2ull * d : d;
            for (uint64_t p = 1 + s; p < limit; p += s) {
                if (!is_odd_prime(bits, p)) continue;
                const Mont32 M((uint32_t)p);
                if (M.pow2(d) != M.one) continue;
                bool minimal = true;
                for (uint32_t q : qs)
                    if (M.pow2(d / q) == M.one) {
                        minimal = false;
                        break;
                    }
                if (minimal) mine.emplace_back((uint32_t)p, d);
            }
        }
#pragma omp critical
        out.insert(out.end(), mine.begin(), mine.end());
    }
    std::sort(out.begin(), out.end());
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) {
        std::perror(argv[3]);
        return 1;
    }
    for (auto &pr : out) {
        const uint32_t w[2] = {pr.first, pr.second};
        std::fwrite(w, sizeof w, 1, f);
    }
    std::fclose(f);
    std::printf("%zu primes\n", out.size());
    return 0;
}
