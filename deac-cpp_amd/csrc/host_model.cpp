// Host side of the DEAC engine; see host_model.hpp.  Compiled with
// -ffp-contract=off: products and sums round separately unless written as
// std::fma, which is how the reference build (g++ -O3, -ffp-contract=fast)
// evaluates the expressions below.
#include "host_model.hpp"

#include <cmath>
#include <cstring>
#include <mutex>

namespace deac {

std::string &last_message()
{
    static thread_local std::string message;
    return message;
}

// ---------------------------------------------------------------- RNG

uint64_t Xoshiro::next()  // rng.hpp:25-39
{
    const uint64_t out = s[0] + s[3];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
    return out;
}

void seed_state(uint64_t seed, uint64_t out[4])  // rng.hpp:46-59, deac.cpp:2089
{
    // splitmix64 (rng.hpp:10-15); each state word keeps 32 bits of an output.
    uint64_t sm = UINT64_C(1407513600) + seed;
    auto splitmix = [&sm]() {
        uint64_t z = (sm += UINT64_C(0x9E3779B97f4A7C15));
        z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
        z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
        return z ^ (z >> 31);
    };
    const uint64_t a = splitmix();
    out[0] = static_cast<uint32_t>(a);
    out[1] = static_cast<uint32_t>(a >> 32);
    const uint64_t b = splitmix();
    out[2] = static_cast<uint32_t>(b);
    out[3] = static_cast<uint32_t>(b >> 32);
}

Gf2Matrix Gf2Matrix::identity()
{
    Gf2Matrix m;
    std::memset(m.col, 0, sizeof m.col);
    for (int b = 0; b < 256; b++) m.col[b][b >> 6] = UINT64_C(1) << (b & 63);
    return m;
}

Gf2Matrix Gf2Matrix::one_step()
{
    Gf2Matrix m;
    for (int b = 0; b < 256; b++) {
        Xoshiro x{{0, 0, 0, 0}};
        x.s[b >> 6] = UINT64_C(1) << (b & 63);
        x.next();
        std::memcpy(m.col[b], x.s, sizeof x.s);
    }
    return m;
}

void Gf2Matrix::apply(const uint64_t in[4], uint64_t out[4]) const
{
    uint64_t o[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; w++) {
        uint64_t bits = in[w];
        while (bits) {
            const int b = __builtin_ctzll(bits) + 64 * w;
            bits &= bits - 1;
            for (int k = 0; k < 4; k++) o[k] ^= col[b][k];
        }
    }
    std::memcpy(out, o, sizeof o);
}

Gf2Matrix Gf2Matrix::operator*(const Gf2Matrix &rhs) const
{
    Gf2Matrix r;
    for (int b = 0; b < 256; b++) apply(rhs.col[b], r.col[b]);
    return r;
}

Gf2Matrix transition_power(uint64_t n)
{
    // powers[k] = T^(2^k), built once.
    static std::once_flag once;
    static std::vector<Gf2Matrix> powers;
    std::call_once(once, [] {
        powers.resize(64);
        powers[0] = Gf2Matrix::one_step();
        for (int k = 1; k < 64; k++) powers[k] = powers[k - 1] * powers[k - 1];
    });
    Gf2Matrix r = Gf2Matrix::identity();
    for (int k = 0; n; k++, n >>= 1)
        if (n & 1) r = powers[k] * r;
    return r;
}

void matrix_rows(const Gf2Matrix &m, uint64_t *rows)
{
    std::memset(rows, 0, sizeof(uint64_t) * 256 * 4);
    for (int b = 0; b < 256; b++)
        for (int r = 0; r < 256; r++)
            if ((m.col[b][r >> 6] >> (r & 63)) & 1) rows[r * 4 + (b >> 6)] |= UINT64_C(1) << (b & 63);
}

void nibble_table(const Gf2Matrix &m, uint64_t *tab)
{
    for (int n = 0; n < 64; n++)
        for (int v = 0; v < 16; v++) {
            uint64_t *e = tab + (static_cast<size_t>(n) * 16 + v) * 4;
            e[0] = e[1] = e[2] = e[3] = 0;
            for (int q = 0; q < 4; q++)
                if ((v >> q) & 1)
                    for (int k = 0; k < 4; k++) e[k] ^= m.col[4 * n + q][k];
        }
}

// ---------------------------------------------------------------- reductions
// matrix_multiply_MxN_by_Nx1 (deac.cpp:102-108) and squared_mean
// (deac.cpp:171-177) are inlined into deac() by the reference build and
// vectorised as in-order reductions: 4-wide blocks of rounded products added
// one by one, then a 2-wide block, then a fused last odd element
// (squared_mean: up to three fused scalar elements).

double dot_fold4(double acc, const double *a, const double *b, size_t n)
{
    size_t i = 0;
    if (n >= 4)
        for (; i + 4 <= n; i += 4)
            for (size_t l = 0; l < 4; l++) acc += a[i + l] * b[i + l];
    if (n - i >= 2) {
        acc += a[i] * b[i];
        acc += a[i + 1] * b[i + 1];
        i += 2;
    }
    if (n - i == 1) acc = std::fma(a[i], b[i], acc);
    return acc;
}

double mean_of(const double *a, size_t n)  // deac.cpp:163-169
{
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i];
    return s / static_cast<double>(n);
}

double squared_mean_of(const double *a, size_t n)
{
    double s = 0.0;
    size_t i = 0;
    for (; i + 4 <= n; i += 4)
        for (size_t l = 0; l < 4; l++) s += a[i + l] * a[i + l];
    for (; i < n; i++) s = std::fma(a[i], a[i], s);
    return s / static_cast<double>(n);
}

// ---------------------------------------------------------------- validation

deac_status validate(const deac_problem &p, std::string &msg)
{
    const bool single_particle = p.spectra_type == DEAC_SPBSF || p.spectra_type == DEAC_SPFSF;
    if (p.spectra_type < DEAC_SPBSF || p.spectra_type > DEAC_BDSF) {
        msg = "Please choose spectra_type from the following options: spbsf, spfsf, bfull, ffull, bdsf";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (p.model < DEAC_MODEL_STANDARD || p.model > DEAC_MODEL_NORMALIZATION) {
        msg = "Unknown deac-cpp model, please select from 'standard, hyperbolic, normalization'";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (!p.imaginary_time || !p.isf || !p.isf_error || (p.genome_size && !p.frequency)) {
        msg = "ISF input file must contain tau, isf, and error arrays of equal length";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (p.number_of_timeslices < 2) {  // deac.cpp:2067-2069
        msg = "ISF input file must contain at least two timeslices";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < p.number_of_timeslices; i++) {  // deac.cpp:2074-2087
        if (!std::isfinite(p.imaginary_time[i]) || !std::isfinite(p.isf[i]) || !std::isfinite(p.isf_error[i])) {
            msg = "ISF input file contains non-finite values";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
        if (single_particle && p.isf[i] > 0.0) {
            msg = "positive ISF values are not supported for single-particle spectra";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
        if (p.isf_error[i] <= 0.0) {
            msg = "ISF input errors must be positive";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
    }
    if (!p.zero_temperature && (!std::isfinite(p.temperature) || p.temperature <= 0.0)) {  // deac.cpp:2094-2098
        msg = "temperature must be positive for non-ZeroT builds";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (p.population_size < 4) {  // deac.cpp:2104-2106
        msg = "population_size must be at least 4";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (p.genome_size < 2) {  // deac.cpp:2126-2128
        msg = "genome_size must be at least 2";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < p.genome_size; i++) {  // deac.cpp:2129-2136
        if (!std::isfinite(p.frequency[i]) || p.frequency[i] < 0.0) {
            msg = "frequencies must be finite and non-negative";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
        if (i > 0 && p.frequency[i] < p.frequency[i - 1]) {
            msg = "frequencies must be sorted in non-decreasing order";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
    }
    if (!std::isfinite(p.first_moment)) {  // deac.cpp:2143-2151
        msg = "first_moment must be finite";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (!std::isfinite(p.third_moment)) {
        msg = "third_moment must be finite";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (p.third_moment >= 0.0 && (!std::isfinite(p.third_moment_error) || p.third_moment_error <= 0.0)) {
        msg = "third_moment_error must be positive when third_moment is used";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    return DEAC_OK;
}

// ---------------------------------------------------------------- model

static double trapezoid_weight(const std::vector<double> &f, size_t j)  // deac.cpp:352-358
{
    const size_t G = f.size();
    if (j == 0) return 0.5 * (f[1] - f[0]);
    if (j == G - 1) return 0.5 * (f[j] - f[j - 1]);
    return 0.5 * (f[j + 1] - f[j - 1]);
}

Model build_model(const deac_problem &p)
{
    Model m;
    m.T = p.number_of_timeslices;
    m.G = p.genome_size;
    m.P = p.population_size;
    m.zero_t = p.zero_temperature != 0;
    m.bdsf = p.spectra_type == DEAC_BDSF;
    m.nsign = (m.zero_t || m.bdsf) ? 1 : 2;
    m.Kc = static_cast<size_t>(m.nsign) * m.G;
    m.model = p.model;
    m.spectra = p.spectra_type;
    m.allow_negative = p.allow_negative_spectral_weight != 0;
    m.renormalize = p.renormalize_each_generation != 0;
    m.normalize = p.normalize != 0;
    m.first_moment = p.first_moment;
    m.third_moment = p.third_moment;
    m.third_moment_error = p.third_moment_error;
    m.use_first = p.first_moment >= 0.0;
    m.use_third = p.third_moment >= 0.0;
    m.temperature = p.temperature;
    m.beta = 1.0 / p.temperature;
    m.cr0 = p.crossover_probability;
    m.tau_cr = p.self_adapting_crossover_probability;
    m.f0 = p.differential_weight;
    m.tau_f = p.self_adapting_differential_weight_probability;
    m.tau.assign(p.imaginary_time, p.imaginary_time + m.T);
    m.isf.assign(p.isf, p.isf + m.T);
    m.err.assign(p.isf_error, p.isf_error + m.T);
    m.freq.assign(p.frequency, p.frequency + m.G);
    m.zeroth_moment = m.isf[0];  // deac.cpp:306

    const double per = (m.spectra == DEAC_SPFSF || m.spectra == DEAC_FFULL) ? -1.0 : 1.0;  // deac.cpp:238-247
    const double beta = m.beta;
    const size_t G = m.G, Kc = m.Kc;

    // Kernel matrix, deac.cpp:318-401.
    m.kernel.assign(m.T * Kc, 0.0);
    for (size_t i = 0; i < m.T; i++) {
        const double t = m.tau[i];
        double *kp = &m.kernel[i * Kc];
        double *km = m.nsign == 2 ? kp + G : nullptr;
        for (size_t j = 0; j < G; j++) {
            const double w = m.freq[j];
            const double df = trapezoid_weight(m.freq, j);
            if (m.zero_t) {  // deac.cpp:392-399 (positive side)
                kp[j] = df * std::exp(-t * w);
                continue;
            }
            if (m.bdsf) {  // detailed balance, deac.cpp:322, 330, 337, 362, 370, 378
                if (m.model == DEAC_MODEL_HYPERBOLIC) {
                    const double bo2mt = 0.5 * beta - t;
                    kp[j] = df * std::cosh(bo2mt * w);
                } else {
                    const double bmt = beta - t;
                    const double num = std::exp(-bmt * w) + std::exp(-t * w);
                    kp[j] = m.model == DEAC_MODEL_NORMALIZATION ? df * (num / (1.0 + std::exp(-beta * w))) : df * num;
                }
                continue;
            }
            if (m.model == DEAC_MODEL_HYPERBOLIC) {  // deac.cpp:323-324, 363-364
                // -0.5*beta and 1.5*beta are rounded before the subtraction
                // (loop invariants in the reference build, never fused).
                const double bo2mtp = -0.5 * beta - t;
                const double bo2mtn = 1.5 * beta - t;
                kp[j] = per * df * std::exp(bo2mtp * w) / 2;
                km[j] = per * df * std::exp(-bo2mtn * w) / 2;
            } else if (m.model == DEAC_MODEL_NORMALIZATION) {  // deac.cpp:376-384
                const double tmb = t - beta;
                const double num = std::exp(-t * w);
                const double den = 1.0 + std::exp(-beta * w);
                kp[j] = df * (num / den);
                const double e = std::exp(tmb * w);
                const double den_n = 1.0 + std::exp(-beta * w);
                km[j] = per * df * (e - e / den_n);
            } else {  // deac.cpp:367-370
                const double tmb = t - beta;
                kp[j] = df * std::exp(-t * w);
                km[j] = per * df * std::exp(tmb * w);
            }
        }
    }

    // Normalization / first-moment / third-moment vectors (deac.cpp:495-538,
    // 628-668, 728-768).  which: 0 = normalization, 1 = first, 3 = third.
    auto terms = [&](int which, std::vector<double> &out) {
        out.assign(Kc, 0.0);
        for (size_t j = 0; j < G; j++) {
            const double w = m.freq[j];
            const double df = trapezoid_weight(m.freq, j);
            double a, b = 0.0;
            if (m.zero_t) {
                a = which == 0 ? df : which == 1 ? df * w : df * std::pow(w, 3) * w;
            } else if (m.bdsf) {  // deac.cpp:508-526, 645-661, 745-761
                double c;  // the model's detailed-balance factor
                if (m.model == DEAC_MODEL_HYPERBOLIC)
                    c = which == 0 ? std::cosh(0.5 * beta * w) : std::sinh(0.5 * beta * w);
                else if (m.model == DEAC_MODEL_NORMALIZATION)
                    c = which == 0 ? 1.0 : std::tanh(0.5 * beta * w);
                else
                    c = which == 0 ? 1.0 + std::exp(-beta * w) : 1.0 - std::exp(-beta * w);
                if (which == 0) a = m.model == DEAC_MODEL_NORMALIZATION ? df : df * c;
                else if (which == 1) a = df * w * c;
                else a = df * std::pow(w, 3) * w * c;
            } else if (m.model == DEAC_MODEL_HYPERBOLIC) {
                const double e = std::exp(-0.5 * beta * w);
                if (which == 0) { a = 0.5 * df / e; b = 0.5 * df * e; }
                else if (which == 1) { a = 0.5 * df * w / e; b = -0.5 * df * w * e; }
                else { a = 0.5 * df * std::pow(w, 3) * w / e; b = -0.5 * df * std::pow(w, 3) * w * e; }
            } else if (m.model == DEAC_MODEL_NORMALIZATION) {
                const double l = 1.0 / (1.0 + std::exp(-beta * w));
                const double e = std::exp(-beta * w);
                const double r = e / (1.0 + e);
                if (which == 0) { a = df * l; b = df * r; }
                else if (which == 1) { a = df * w * l; b = -df * w * r; }
                else { a = df * std::pow(w, 3) * w * l; b = -df * std::pow(w, 3) * w * r; }
            } else {
                if (which == 0) { a = df; b = df; }
                else if (which == 1) { a = df * w; b = -df * w; }
                else { a = df * std::pow(w, 3) * w; b = -df * std::pow(w, 3) * w; }
            }
            out[j] = a;
            if (m.nsign == 2) out[G + j] = b;
        }
    };
    if (m.normalize || m.renormalize) terms(0, m.nterm);
    if (m.use_first) terms(1, m.fterm);
    if (m.use_third) terms(3, m.tterm);

    // The -1st-moment term of the detailed-balance build (deac.cpp:875-894;
    // the zero-temperature build turns it off, deac.cpp:230-236): trapezoid
    // weights over tau, the ISF's moment summed with fused products, and an
    // error that keeps the last timeslice's term only (the reference assigns
    // rather than accumulates it).
    m.use_nfm = m.bdsf && !m.zero_t && p.use_negative_first_moment != 0;
    if (m.use_nfm) {
        m.nfm_dt.assign(m.T, 0.0);
        double v = 0.0, e = 0.0;
        for (size_t j = 0; j < m.T; j++) {
            const double d = trapezoid_weight(m.tau, j);
            m.nfm_dt[j] = d;
            v = std::fma(m.isf[j], d, v);
            e = (m.err[j] * m.err[j]) * (d * d);
        }
        m.nfm_value = v;
        m.nfm_error = std::sqrt(e);
    }
    return m;
}

void initial_population(const Model &m, Xoshiro &rng, std::vector<double> &pop)
{
    const size_t Kc = m.Kc, G = m.G;
    pop.assign(m.P * Kc, 0.0);
    for (size_t i = 0; i < m.P; i++)  // deac.cpp:430-450: + and - draws interleaved per gene
        for (size_t j = 0; j < G; j++)
            for (int s = 0; s < m.nsign; s++) {
                double u = rng.uniform();
                if (m.allow_negative) u -= 0.5;
                pop[i * Kc + s * G + j] = u;
            }
    if (m.nsign == 2)
        for (size_t i = 0; i < m.P; i++) pop[i * Kc + G] = pop[i * Kc];  // deac.cpp:451-458
    if (m.normalize)  // deac.cpp:584-602
        for (size_t i = 0; i < m.P; i++) {
            double *x = &pop[i * Kc];
            double norm = 0.0;
            for (int s = 0; s < m.nsign; s++) norm = dot_fold4(norm, x + s * G, &m.nterm[s * G], G);
            for (size_t k = 0; k < Kc; k++) x[k] *= m.zeroth_moment / norm;
        }
}

void output_transform(const Model &m, const double *xb, double *dsf, double *freq)
{
    const size_t G = m.G;
    const double beta = m.beta;
    for (size_t i = 0; i < G; i++) {
        const double w = m.freq[i];
        if (m.nsign == 1) {  // zero temperature or bdsf: one sign (deac.cpp:1605-1671)
            double v = xb[i];
            if (!m.zero_t && m.model == DEAC_MODEL_HYPERBOLIC) v = 0.5 * xb[i] * std::exp(0.5 * beta * w);
            else if (!m.zero_t && m.model == DEAC_MODEL_NORMALIZATION) v = xb[i] / (1.0 + std::exp(-beta * w));
            freq[i] = w;
            dsf[i] = v;
            continue;
        }
        const size_t ip = G + i - 1, in = G - i - 1;  // deac.cpp:1600-1608
        freq[ip] = w;
        freq[in] = -w;
        const double sp = xb[i], sn = xb[G + i];
        double vp, vn;
        if (m.model == DEAC_MODEL_HYPERBOLIC) {  // deac.cpp:1620-1635
            if (m.spectra == DEAC_SPBSF) {
                vp = -2.0 * M_PI * sp * std::sinh(beta * w / 2);
                vn = 2.0 * M_PI * sn * std::sinh(beta * w / 2);
            } else if (m.spectra == DEAC_SPFSF) {
                vp = -2.0 * M_PI * sp * std::cosh(beta * w / 2);
                vn = -2.0 * M_PI * sn * std::cosh(beta * w / 2);
            } else {
                vp = 0.5 * sp * std::exp(0.5 * beta * w);
                vn = 0.5 * sn * std::exp(-0.5 * beta * w);
            }
        } else if (m.model == DEAC_MODEL_NORMALIZATION) {  // deac.cpp:1653-1667
            if (m.spectra == DEAC_SPBSF) {
                vp = -2.0 * M_PI * sp * std::tanh(0.5 * beta * w);
                vn = 2.0 * M_PI * sn * std::tanh(0.5 * beta * w);
            } else if (m.spectra == DEAC_SPFSF) {
                vp = -2.0 * M_PI * sp;
                vn = -2.0 * M_PI * sn;
            } else {
                vp = sp / (1.0 + std::exp(-beta * w));
                vn = sn / (1.0 + std::exp(beta * w));
            }
        } else {  // deac.cpp:1637-1651
            if (m.spectra == DEAC_SPBSF) {
                vp = -2.0 * M_PI * sp * (1.0 - std::exp(-beta * w));
                vn = -2.0 * M_PI * sn * (1.0 - std::exp(beta * w));
            } else if (m.spectra == DEAC_SPFSF) {
                vp = -2.0 * M_PI * sp * (1.0 + std::exp(-beta * w));
                vn = -2.0 * M_PI * sn * (1.0 + std::exp(beta * w));
            } else {
                vp = sp;
                vn = sn;
            }
        }
        dsf[ip] = vp;
        dsf[in] = vn;
    }
}

}  // namespace deac
