// ctd_sequences.hpp -- the built-in sequence families behind the plugin API.
//
// Each family is a pure math sequence in the sense of number_sequence.hpp:19-53
// (host-side residue functions, exact 128-bit arithmetic) that also names the
// device kernel propagating it (device_seq_id, device_param()).  The wrappers
// with short names and the dispatch follow num_seq_dispatch.hpp:119-168.
// Families: mersenne.hpp, riesel_proth.hpp, riesel_proth_fixed_n.hpp,
// smarandache.hpp, fibonacci.hpp, lucas.hpp, tribonacci.hpp (all under
// number_sequences/).
#pragma once

#include <charconv>
#include <cstdint>
#include <ostream>
#include <string>
#include <utility>
#include <vector>

namespace cutrialdive {

using u128 = unsigned __int128;

namespace detail {
inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t d) { return (uint64_t)(((u128)a * b) % d); }
inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t d) { return a < d - b ? a + b : a - (d - b); }
inline uint64_t submod(uint64_t a, uint64_t b, uint64_t d) { return a >= b ? a - b : d - (b - a); }
inline uint64_t powmod(uint64_t a, uint64_t e, uint64_t d)
{
    uint64_t r = 1 % d, m = a % d;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod(r, m, d);
        m = mulmod(m, m, d);
    }
    return r;
}
inline uint64_t digits(uint64_t b, uint64_t n)
{
    uint64_t k = 1;
    while ((n /= b)) ++k;
    return k;
}
} // namespace detail

// Unsigned big integer (little-endian 64-bit limbs, no leading zero limb):
// the value() / print_value() part of the plugin API.  Only the few
// operations the built-in families need; speed is not a concern here.
struct big_uint {
    std::vector<uint64_t> limbs;
    big_uint() = default;
    big_uint(uint64_t v)
    {
        if (v) limbs.push_back(v);
    }
    void trim()
    {
        while (!limbs.empty() && !limbs.back()) limbs.pop_back();
    }
    uint64_t mod(uint64_t d) const
    {
        u128 r = 0;
        for (auto it = limbs.rbegin(); it != limbs.rend(); ++it) r = ((r << 64) | *it) % d;
        return (uint64_t)r;
    }
    // this = this * m + a
    void mul_add(uint64_t m, uint64_t a)
    {
        u128 c = a;
        for (auto &l : limbs) {
            u128 t = (u128)l * m + c;
            l = (uint64_t)t;
            c = t >> 64;
        }
        if (c) limbs.push_back((uint64_t)c);
        trim();
    }
    big_uint &operator+=(big_uint const &o)
    {
        if (limbs.size() < o.limbs.size()) limbs.resize(o.limbs.size(), 0);
        u128 c = 0;
        for (std::size_t i = 0; i < limbs.size(); ++i) {
            u128 t = (u128)limbs[i] + (i < o.limbs.size() ? o.limbs[i] : 0) + c;
            limbs[i] = (uint64_t)t;
            c = t >> 64;
        }
        if (c) limbs.push_back((uint64_t)c);
        return *this;
    }
    // this -= a, requires this >= a
    void sub_small(uint64_t a)
    {
        for (std::size_t i = 0; a && i < limbs.size(); ++i) {
            uint64_t l = limbs[i];
            limbs[i] = l - a;
            a = l < a;
        }
        trim();
    }
    void shl(uint64_t bits)
    {
        if (limbs.empty()) return;
        uint64_t w = bits / 64, b = bits % 64;
        if (b) {
            uint64_t carry = 0;
            for (auto &l : limbs) {
                uint64_t nl = (l << b) | carry;
                carry = l >> (64 - b);
                l = nl;
            }
            if (carry) limbs.push_back(carry);
        }
        limbs.insert(limbs.begin(), w, 0);
    }
    std::string to_string() const
    {
        if (limbs.empty()) return "0";
        constexpr uint64_t chunk = 10000000000000000000ull; // 10^19
        std::vector<uint64_t> q = limbs;
        std::vector<uint64_t> parts;
        while (!q.empty()) {
            u128 r = 0;
            for (auto it = q.rbegin(); it != q.rend(); ++it) {
                u128 cur = (r << 64) | *it;
                *it = (uint64_t)(cur / chunk);
                r = cur % chunk;
            }
            parts.push_back((uint64_t)r);
            while (!q.empty() && !q.back()) q.pop_back();
        }
        std::string s = std::to_string(parts.back());
        for (auto it = parts.rbegin() + 1; it != parts.rend(); ++it) {
            std::string d = std::to_string(*it);
            s += std::string(19 - d.size(), '0') + d;
        }
        return s;
    }
    bool operator==(big_uint const &) const = default;
};
inline std::ostream &operator<<(std::ostream &out, big_uint const &v) { return out << v.to_string(); }

inline std::string u128_to_string(u128 v)
{
    if (!v) return "0";
    std::string s;
    for (; v; v /= 10) s.insert(s.begin(), char('0' + int(v % 10)));
    return s;
}

// k * 2^e + sign (sign in {-1, 0, +1}); k * 2^e >= 1 when sign is -1
inline big_uint k_pow2_pm(uint64_t k, uint64_t e, int sign)
{
    big_uint v{k};
    v.shl(e);
    if (sign > 0) v += big_uint{1};
    if (sign < 0) v.sub_small(1);
    return v;
}

// ---- Mersenne M(n) = 2^n - 1 (mersenne.hpp:19-124)
struct mersenne {
    using index_type = uint64_t;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    static constexpr bool initialize_from_value = false;
    static constexpr uint32_t device_seq_id = CTD_SEQ_MERSENNE;
    uint64_t device_param() const { return 0; }
    static char const *short_name() { return "M"; }
    static residue_type value_mod_2(index_type n) { return n > 0; }
    static residue_type value_mod_mu(index_type n, residue_type d, mu_both_t)
    {
        if (n <= 1) return (n == 0 || d == 1) ? 0 : 1;
        uint64_t r = detail::powmod(2, n, d);
        return r ? r - 1 : d - 1;
    }
    static residue_type next_value_mod(residue_type r, index_type, residue_type d)
    {
        return detail::addmod(detail::addmod(r, r, d), 1 % d, d);
    }
    static residue_type next_value_mod_2(residue_type, index_type) { return 1; }
    static big_uint value(index_type n) { return n ? k_pow2_pm(1, n, -1) : big_uint{}; }
    static std::ostream &print_value(index_type n, std::ostream &out) { return out << value(n); }
    static std::ostream &print_expression(index_type n, std::ostream &out) { return out << "2^" << n << "-1"; }
};

// ---- Proth / Riesel k 2^n +- 1 (riesel_proth.hpp:23-179)
enum class plus_or_minus { plus, minus };

template <plus_or_minus PM>
class riesel_proth_math {
public:
    using index_type = uint64_t;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    static constexpr uint32_t device_seq_id = PM == plus_or_minus::plus ? CTD_SEQ_PROTH : CTD_SEQ_RIESEL;
    explicit riesel_proth_math(uint64_t k) : k_(k)
    {
        if (!k)
            throw std::invalid_argument(std::string{PM == plus_or_minus::plus ? "Proth" : "Riesel"} +
                                        " parameter k must be non-zero");
    }
    uint64_t get_k() const { return k_; }
    uint64_t device_param() const { return k_; }
    residue_type value_mod_2(index_type n) const { return n > 0 ? 1 : ((k_ & 1) ^ 1); }
    residue_type value_mod_mu(index_type n, residue_type d, mu_both_t) const
    {
        uint64_t r = detail::mulmod(detail::powmod(2, n, d), k_ % d, d);
        return PM == plus_or_minus::plus ? detail::addmod(r, 1 % d, d) : detail::submod(r, 1 % d, d);
    }
    residue_type next_value_mod_2(residue_type, index_type) const { return 1; }
    residue_type next_value_mod(residue_type r, index_type, residue_type d) const
    {
        r = detail::addmod(r, r, d);
        return PM == plus_or_minus::plus ? detail::submod(r, 1 % d, d) : detail::addmod(r, 1 % d, d);
    }
    big_uint value(index_type n) const { return k_pow2_pm(k_, n, PM == plus_or_minus::plus ? 1 : -1); }
    std::ostream &print_value(index_type n, std::ostream &out) const { return out << value(n); }
    std::ostream &print_expression(index_type n, std::ostream &out) const
    {
        return out << k_ << "*2^" << n << " " << (PM == plus_or_minus::plus ? "+" : "-") << " 1";
    }

private:
    uint64_t k_;
};

template <plus_or_minus PM>
class riesel_proth {
public:
    riesel_proth(riesel_proth_math<PM> const &m)
        : m_(m), name_(std::string{PM == plus_or_minus::plus ? "Pr_" : "R_"} + std::to_string(m.get_k()))
    {
    }
    riesel_proth(uint64_t k) : riesel_proth(riesel_proth_math<PM>{k}) {}
    riesel_proth_math<PM> math_sequence() const { return m_; }
    char const *short_name() const { return name_.c_str(); }

private:
    riesel_proth_math<PM> m_;
    std::string name_;
};
using proth = riesel_proth<plus_or_minus::plus>;
using riesel = riesel_proth<plus_or_minus::minus>;

// ---- fixed-n Proth / Riesel k 2^n +- 1 indexed by k (riesel_proth_fixed_n.hpp:20-181)
template <plus_or_minus PM>
class riesel_proth_fixed_n_math {
public:
    using index_type = uint64_t;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    using state_type = uint64_t;
    static const bool has_state = true;
    static constexpr uint32_t device_seq_id =
        PM == plus_or_minus::plus ? CTD_SEQ_PROTH_FIXED_N : CTD_SEQ_RIESEL_FIXED_N;
    explicit riesel_proth_fixed_n_math(uint64_t n) : n_(n)
    {
        if (!n) throw std::invalid_argument{"Invalid n value zero for Riesel/Proth fixed n"};
    }
    uint64_t get_n() const { return n_; }
    uint64_t device_param() const { return n_; }
    state_type make_state(index_type, residue_type d, mu_both_t) const { return detail::powmod(2, n_, d); }
    residue_type value_mod_2(index_type) const { return 1; }
    residue_type value_mod_mu(index_type k, residue_type d, mu_both_t, state_type &st) const
    {
        uint64_t r = detail::mulmod(k % d, st, d);
        return PM == plus_or_minus::plus ? detail::addmod(r, 1 % d, d) : detail::submod(r, 1 % d, d);
    }
    residue_type next_value_mod_2(residue_type, index_type) const { return 1; }
    residue_type next_value_mod(residue_type r, index_type, residue_type d, state_type &st) const
    {
        return detail::addmod(r, st, d);
    }
    big_uint value(index_type k) const
    {
        if (!k && PM == plus_or_minus::minus) throw std::domain_error{"0*2^n-1 is negative"};
        return k_pow2_pm(k, n_, PM == plus_or_minus::plus ? 1 : -1);
    }
    std::ostream &print_value(index_type k, std::ostream &out) const { return out << value(k); }
    std::ostream &print_expression(index_type k, std::ostream &out) const
    {
        return out << k << "*2^" << n_ << (PM == plus_or_minus::plus ? '+' : '-') << "1";
    }

private:
    uint64_t n_;
};

template <plus_or_minus PM>
class riesel_proth_fixed_n {
public:
    riesel_proth_fixed_n(riesel_proth_fixed_n_math<PM> const &m)
        : m_(m), name_(std::string{PM == plus_or_minus::plus ? "Prn_" : "Rn_"} + std::to_string(m.get_n()))
    {
    }
    riesel_proth_fixed_n(uint64_t n) : riesel_proth_fixed_n(riesel_proth_fixed_n_math<PM>{n}) {}
    riesel_proth_fixed_n_math<PM> math_sequence() const { return m_; }
    char const *short_name() const { return name_.c_str(); }

private:
    riesel_proth_fixed_n_math<PM> m_;
    std::string name_;
};
using proth_fixed_n = riesel_proth_fixed_n<plus_or_minus::plus>;
using riesel_fixed_n = riesel_proth_fixed_n<plus_or_minus::minus>;

// ---- Smarandache base b: concatenation of 1..n in base b (smarandache.hpp:20-207)
template <uint64_t Base>
class smarandache {
    static_assert(Base > 1, "Base must be at least 2");

public:
    using index_type = uint64_t;
    using value_type = big_uint;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    static constexpr bool initialize_from_value = true;
    static constexpr uint64_t base = Base;
    static constexpr uint32_t device_seq_id = CTD_SEQ_SMARANDACHE;
    uint64_t device_param() const { return Base; }
    static char const *short_name()
    {
        static const std::string name = Base == 10 ? std::string{"Sm"} : "Sm" + std::to_string(Base);
        return name.c_str();
    }
    static value_type value(index_type n)
    {
        big_uint v;
        for (uint64_t i = 1; i <= n; ++i) {
            uint64_t m = 1;
            for (uint64_t k = detail::digits(Base, i); k; --k) m *= Base;
            v.mul_add(m, i);
        }
        return v;
    }
    static residue_type value_mod_2(index_type n)
    {
        if constexpr (!(Base & 1)) return n & 1;
        else return (n & 3) && ((n & 3) != 3);
    }
    static residue_type next_value_mod_2(residue_type r, index_type n)
    {
        if constexpr (!(Base & 1)) return (n & 1) ^ 1;
        else return r ^ (n & 1) ^ 1;
    }
    static residue_type next_value_mod(residue_type r, index_type n, residue_type p)
    {
        u128 m = 1;
        for (uint64_t k = detail::digits(Base, n + 1); k; --k) m = (m * Base) % p;
        return (uint64_t)(((u128)r * (uint64_t)m + (n + 1) % p) % p);
    }
    static residue_type next_value_mod_mu(residue_type r, index_type n, residue_type p, mu_type)
    {
        return next_value_mod(r, n, p);
    }
    static std::ostream &print_value(index_type n, std::ostream &out) { return out << value(n); }

    // Closed form of smarandache.cpp:160-181: with k = floor(log_b n) and
    // B = b^(k+1),  Sm_b(n) = (C_k * b^((k+1)(n - b^k + 1)) - (B - 1) n - B) / (B - 1)^2,
    // where C_0 = b and C_i is written out from C_{i-1} (smarandache.cpp:55-77).
    static std::ostream &print_expression(index_type n, std::ostream &out)
    {
        if (n < 2) return out << n;
        uint64_t k = 0;
        for (uint64_t t = n; t /= Base;) ++k;
        const std::string b = std::to_string(Base);
        auto power = [&](uint64_t e) { return b + "^" + std::to_string(e); };
        auto has_op = [](std::string const &x) {
            for (char c : x)
                if (c < '0' || c > '9') return true;
            return false;
        };
        std::string c = b;
        u128 b_pow = 1; // b^(i-1)
        for (uint64_t i = 1; i <= k; ++i) {
            const std::string bi = i != 1 ? power(i) : b;
            const std::string pred = has_op(c) ? "(" + c + ")" : c;
            c = "(" + pred + "*" + b + "^" + u128_to_string((u128)(Base - 1) * i * b_pow) + " - (" + bi +
                " - 1)^2 - " + bi + ")/(" + bi + " - 1)^2*(" + power(i + 1) + " - 1)^2 + " + power(2 * i + 1) +
                " - " + bi + " + 1";
            b_pow *= Base;
        }
        if (has_op(c)) c = "(" + c + ")";
        const std::string bk1 = k == 0 ? b : power(k + 1);
        u128 bk = 1;
        for (uint64_t i = 0; i < k; ++i) bk *= Base;
        out << "(" << c << "*" << Base << "^" << u128_to_string((u128)(k + 1) * (n - (uint64_t)bk + 1)) << " - ("
            << bk1 << " - 1)*" << n << " - " << bk1 << ")/(" << bk1 << " - 1)^2";
        return out;
    }
};

// ---- Fibonacci (fibonacci.hpp:18-104); F(0) = 0
class fibonacci {
public:
    struct state {
        uint64_t previous;
        uint64_t current;
    };
    using index_type = uint64_t;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    using state_type = state;
    static const bool has_state = true;
    static constexpr uint32_t device_seq_id = CTD_SEQ_FIBONACCI;
    uint64_t device_param() const { return 0; }
    static char const *short_name() { return "I"; }
    // {F(n-1), F(n)}; n = 0 gives {0, 0} exactly as fibonacci.hpp:68-71
    state_type make_state(index_type n, residue_type d, mu_both_t) const
    {
        if (n <= 1) return {0, n};
        uint64_t a = 0, b = 1 % d; // (F(k), F(k+1)) with k = 0
        const uint64_t m = n - 1;
        for (int bit = 63 - __builtin_clzll(m); bit >= 0; --bit) {
            uint64_t c = detail::mulmod(a, detail::submod(detail::addmod(b, b, d), a, d), d);
            uint64_t e = detail::addmod(detail::mulmod(a, a, d), detail::mulmod(b, b, d), d);
            if ((m >> bit) & 1) {
                a = e;
                b = detail::addmod(c, e, d);
            } else {
                a = c;
                b = e;
            }
        }
        return {a, b};
    }
    residue_type value_mod_2(index_type n) const { return n % 3 == 0 ? 0 : 1; }
    residue_type next_value_mod_2(residue_type, index_type n) const { return (n + 1) % 3 == 0 ? 0 : 1; }
    residue_type value_mod(index_type, residue_type, state_type &st) const { return st.current; }
    residue_type next_value_mod(residue_type r, index_type, residue_type d, state_type &st) const
    {
        uint64_t t = detail::addmod(r, st.previous, d);
        st.previous = r;
        return t;
    }
    // G(0) = a, G(1) = b, G(n+1) = G(n) + G(n-1)
    static big_uint linear2(uint64_t a, uint64_t b, index_type n)
    {
        big_uint x{a}, y{b};
        if (!n) return x;
        for (index_type i = 1; i < n; ++i) {
            x += y;
            std::swap(x, y);
        }
        return y;
    }
    big_uint value(index_type n) const { return linear2(0, 1, n); }
    std::ostream &print_value(index_type n, std::ostream &out) const { return out << value(n); }
};

// ---- Lucas (lucas.hpp:15-79); L(0) = 2, started from the Fibonacci state
class lucas {
public:
    using index_type = uint64_t;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    using state_type = fibonacci::state;
    static const bool has_state = true;
    static constexpr uint32_t device_seq_id = CTD_SEQ_LUCAS;
    uint64_t device_param() const { return 0; }
    static char const *short_name() { return "L"; }
    state_type make_state(index_type n, residue_type d, mu_both_t mu) const
    {
        auto f = fibonacci{}.make_state(n, d, mu);
        return {detail::submod(detail::addmod(f.current, f.current, d), f.previous, d),
                detail::addmod(f.current, detail::addmod(f.previous, f.previous, d), d)};
    }
    residue_type value_mod_2(index_type n) const { return n % 3 == 0 ? 0 : 1; }
    residue_type next_value_mod_2(residue_type, index_type n) const { return (n + 1) % 3 == 0 ? 0 : 1; }
    residue_type value_mod(index_type, residue_type, state_type &st) const { return st.current; }
    residue_type next_value_mod(residue_type r, index_type n, residue_type d, state_type &st) const
    {
        return fibonacci{}.next_value_mod(r, n, d, st);
    }
    big_uint value(index_type n) const { return fibonacci::linear2(2, 1, n); }
    std::ostream &print_value(index_type n, std::ostream &out) const { return out << value(n); }
};

// ---- Tribonacci T(0) = T(1) = 0, T(2) = 1 (tribonacci.hpp:19-143)
class tribonacci {
public:
    struct state {
        uint64_t t_n, t_n1, t_n2;
    };
    using index_type = uint64_t;
    using residue_type = uint64_t;
    using mu_type = mu_both_t;
    using state_type = state;
    static const bool has_state = true;
    static constexpr uint32_t device_seq_id = CTD_SEQ_TRIBONACCI;
    uint64_t device_param() const { return 0; }
    static char const *short_name() { return "T"; }
    state_type make_state(index_type n, residue_type d, mu_both_t) const
    {
        // 3x3 companion matrix power: (T(n), T(n+1), T(n+2)) = M^n (0, 0, 1)
        uint64_t M[3][3] = {{0, 1, 0}, {0, 0, 1}, {1 % d, 1 % d, 1 % d}};
        uint64_t R[3][3] = {{1 % d, 0, 0}, {0, 1 % d, 0}, {0, 0, 1 % d}};
        auto mul = [d](uint64_t (&A)[3][3], uint64_t const (&B)[3][3]) {
            uint64_t C[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    uint64_t s = 0;
                    for (int k = 0; k < 3; ++k) s = detail::addmod(s, detail::mulmod(A[i][k], B[k][j], d), d);
                    C[i][j] = s;
                }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) A[i][j] = C[i][j];
        };
        for (uint64_t e = n; e; e >>= 1) {
            if (e & 1) mul(R, M);
            uint64_t T[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) T[i][j] = M[i][j];
            mul(M, T);
        }
        return {R[0][2], R[1][2], R[2][2]};
    }
    residue_type value_mod_2(index_type n) const { return (n & 3) >> 1; }
    residue_type next_value_mod_2(residue_type, index_type n) const { return ((n + 1) & 3) >> 1; }
    residue_type value_mod(index_type, residue_type, state_type &st) const { return st.t_n; }
    residue_type next_value_mod(residue_type, index_type, residue_type d, state_type &st) const
    {
        uint64_t t = detail::addmod(detail::addmod(st.t_n, st.t_n1, d), st.t_n2, d);
        st = {st.t_n1, st.t_n2, t};
        return st.t_n;
    }
    big_uint value(index_type n) const
    {
        big_uint a{0}, b{0}, c{1}; // T(i), T(i+1), T(i+2)
        for (index_type i = 0; i < n; ++i) {
            big_uint t = a;
            t += b;
            t += c;
            a = std::move(b);
            b = std::move(c);
            c = std::move(t);
        }
        return a;
    }
    std::ostream &print_value(index_type n, std::ostream &out) const { return out << value(n); }
};

// ---------------------------------------------------------------- dispatch (num_seq_dispatch.hpp:119-168)

namespace details {
template <typename IntT>
std::optional<IntT> parse_int(std::string_view s)
{
    IntT v{};
    auto [p, ec] = std::from_chars(s.data(), s.data() + s.size(), v);
    if (ec == std::errc()) return v;
    return {};
}

template <typename Func, uint64_t... Bs>
auto dispatch_smarandache_impl(uint64_t base, Func &&f, std::integer_sequence<uint64_t, Bs...>)
{
    using R = decltype(f.template operator()<smarandache<2>>());
    bool found = false;
    if constexpr (std::is_void_v<R>) {
        ((base == Bs + 2 ? (found = true, f.template operator()<smarandache<Bs + 2>>(), 0) : 0), ...);
        if (!found)
            throw std::runtime_error{"Base `" + std::to_string(base) + "' is not supported for Smarandache numbers"};
    } else {
        std::optional<R> r;
        ((base == Bs + 2 ? (found = true, r.emplace(f.template operator()<smarandache<Bs + 2>>()), 0) : 0), ...);
        if (!found)
            throw std::runtime_error{"Base `" + std::to_string(base) + "' is not supported for Smarandache numbers"};
        return std::move(*r);
    }
}
} // namespace details

template <typename IndexT, typename Func>
auto dispatch_num_seq(num_seq_spec spec, Func &&f)
{
    switch (spec.seq_id) {
    case num_seq_id::mersenne: return f.template operator()<mersenne>();
    case num_seq_id::proth:
    case num_seq_id::riesel: {
        auto k = details::parse_int<uint64_t>(spec.seq_params);
        bool pr = spec.seq_id == num_seq_id::proth;
        if (!k)
            throw std::runtime_error{"Missing or invalid k value `" + spec.seq_params + "' in " +
                                     (pr ? "Proth number spec" : "Riesel sequence spec")};
        return pr ? f.template operator()<proth>(*k) : f.template operator()<riesel>(*k);
    }
    case num_seq_id::smarandache: {
        auto b = spec.seq_params.empty() ? std::optional<uint64_t>{10} : details::parse_int<uint64_t>(spec.seq_params);
        if (!b) throw std::runtime_error{"Invalid base `" + spec.seq_params + "' for Smarandache numbers"};
        return details::dispatch_smarandache_impl(*b, f, std::make_integer_sequence<uint64_t, 63>{});
    }
    case num_seq_id::fibonacci: return f.template operator()<fibonacci>();
    case num_seq_id::lucas: return f.template operator()<lucas>();
    case num_seq_id::proth_fixed_n:
    case num_seq_id::riesel_fixed_n: {
        auto n = details::parse_int<uint64_t>(spec.seq_params);
        bool pr = spec.seq_id == num_seq_id::proth_fixed_n;
        if (!n)
            throw std::runtime_error{"Missing or invalid n value `" + spec.seq_params + "' in " +
                                     (pr ? "Proth number spec" : "Riesel sequence spec")};
        return pr ? f.template operator()<proth_fixed_n>(*n) : f.template operator()<riesel_fixed_n>(*n);
    }
    case num_seq_id::tribonacci: return f.template operator()<tribonacci>();
    }
    throw std::runtime_error{"Unexpected number sequence id: " + std::to_string(static_cast<int>(spec.seq_id))};
}

} // namespace cutrialdive
