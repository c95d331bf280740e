// Input parsers (.opt, .mci), derived parameters, tally finalisation and the CSV/.mco
// writers of the drop-in output layout.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "host_common.hpp"

namespace octmps {

static thread_local std::string t_err;

void set_error(const std::string &msg) { t_err = msg; }

octmps_status fail(octmps_status st, const std::string &msg) {
    t_err = msg;
    return st;
}

namespace {

// Line scanners with the reference's semantics (readfloats/readdoubles/readints,
// src/octmps_io.cpp:444-494): skip lines that scan no number (comments, blank lines),
// fail when a line holds more numbers than requested or the file ends.
int read_numbers(FILE *fp, int n, const char *fmt, void *dst0, void *dst1, void *dst2,
                 void *dst3, void *dst4, size_t elem, int buflen) {
    char line[512];
    int ii = 0;
    while (ii <= 0) {
        if (feof(fp)) return 0;
        if (!fgets(line, buflen, fp)) line[0] = 0;
        std::memset(dst0, 0, elem);
        std::memset(dst1, 0, elem);
        std::memset(dst2, 0, elem);
        std::memset(dst3, 0, elem);
        std::memset(dst4, 0, elem);
        ii = sscanf(line, fmt, dst0, dst1, dst2, dst3, dst4);
        if (ii > n) return 0;
    }
    return 1;
}

int readfloats(int n, float *t, FILE *fp) {
    return read_numbers(fp, n, "%f %f %f %f %f", &t[0], &t[1], &t[2], &t[3], &t[4],
                        sizeof(float), 200);
}
int readdoubles(int n, double *t, FILE *fp) {
    return read_numbers(fp, n, "%lf %lf %lf %lf %lf", &t[0], &t[1], &t[2], &t[3], &t[4],
                        sizeof(double), 200);
}
int readints(int n, int *t, FILE *fp) {
    return read_numbers(fp, n, "%d %d %d %d %d", &t[0], &t[1], &t[2], &t[3], &t[4],
                        sizeof(int), 256);
}

bool ischar(char a) { return (a >= 'A' && a <= 'Z') || (a >= 'a' && a <= 'z'); }

struct File {
    FILE *fp;
    explicit File(const char *p, const char *mode) : fp(fopen(p, mode)) {}
    ~File() {
        if (fp) fclose(fp);
    }
};

}  // namespace
}  // namespace octmps

using namespace octmps;

extern "C" const char *octmps_status_string(int s) {
    switch (s) {
        case OCTMPS_OK: return "ok";
        case OCTMPS_E_INVALID: return "invalid argument";
        case OCTMPS_E_IO: return "i/o error";
        case OCTMPS_E_PARSE: return "parse error";
        case OCTMPS_E_PROBE: return "probe not on the tissue surface";
        case OCTMPS_E_RANGE: return "out of range";
        case OCTMPS_E_SEED: return "invalid seed";
        case OCTMPS_E_NOMEM: return "out of memory";
        case OCTMPS_E_DEVICE: return "device error";
        case OCTMPS_E_BUSY: return "busy";
        case OCTMPS_E_CANCELLED: return "cancelled";
        case OCTMPS_E_NOT_FOUND: return "not found";
        case OCTMPS_E_TIMEOUT: return "timeout";
        case OCTMPS_E_FAILED: return "failed";
        default: return "unknown status";
    }
}

extern "C" const char *octmps_last_error(void) { return t_err.c_str(); }

extern "C" const char *octmps_version(void) { return "octmps-mi355x 0.1 (gfx950)"; }

extern "C" double octmps_rspecular(double ni, double nt) {
    // Rspecular, src/main.cu:29-37.
    double temp = (ni - nt) / (ni + nt);
    return temp * temp;
}

extern "C" void octmps_params_default(octmps_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->n_regions = 1;
    p->regions[0].n = 1.0;
    p->type_simulation = 1;
    p->type_bias = 37;
    p->backward_bias_coefficient = 0.925;
    p->target_depth_min = 0.0;
    p->target_depth_max = 0.1;
    p->coherence_length_source = 0.0015;
    p->max_collecting_radius = 10e-4;
    p->max_collecting_angle_deg = 5.0;
    p->probability_additional_bias = 0.5;
    p->max_relative_contribution = 0.01;
    p->rnd_step_size_in_tissue = 0.0;
    p->max_filtered_optical_depth = 0.12;
    p->num_depth_bins = 480;
    p->optical_depth_shift = 0.0;
    p->number_of_photons = 100000;
}

extern "C" octmps_status octmps_read_opt(const char *path, int32_t run_index, octmps_params *p,
                                         int32_t *n_runs) {
    // read_simulation_data, src/octmps_io.cpp:504-645.
    if (!path || !p) return fail(OCTMPS_E_INVALID, "null argument");
    File f(path, "r");
    if (!f.fp) return fail(OCTMPS_E_IO, std::string("cannot open ") + path);
    float ftemp[5];
    double dtemp[5];
    int itemp[5];
    if (!readfloats(1, ftemp, f.fp)) return fail(OCTMPS_E_PARSE, "error reading file version");
    if (!readints(1, itemp, f.fp)) return fail(OCTMPS_E_PARSE, "error reading number of runs");
    int nsim = itemp[0];
    if (n_runs) *n_runs = nsim;
    if (nsim <= 0) return fail(OCTMPS_E_PARSE, "number of runs must be positive");
    if (run_index < 0 || run_index >= nsim) return fail(OCTMPS_E_RANGE, "run index out of range");
    for (int i = 0; i <= run_index; ++i) {
        char line[256], str[256];
        char aorb = 0;
        int ii = 0;
        while (ii <= 0) {
            if (!fgets(line, 256, f.fp)) line[0] = 0;
            ii = sscanf(line, "%255s %c", str, &aorb);
            if (feof(f.fp) || ii > 2) return fail(OCTMPS_E_PARSE, "error reading output filename");
            if (ii > 0) ii = ischar(str[0]);
        }
        unsigned int nph = 0;
        ii = 0;
        while (ii <= 0) {
            if (!fgets(line, 256, f.fp)) line[0] = 0;
            nph = 0;
            ii = sscanf(line, "%u", &nph);
            if (feof(f.fp) || ii > 1) return fail(OCTMPS_E_PARSE, "error reading number of photons");
        }
        if (!readints(1, itemp, f.fp)) return fail(OCTMPS_E_PARSE, "error reading No. of regions");
        int nreg = itemp[0] + 1;
        if (nreg < 2 || nreg > OCTMPS_MAX_REGIONS)
            return fail(OCTMPS_E_RANGE, "number of regions out of range (MAX_REGIONS=100)");
        if (!readdoubles(1, dtemp, f.fp))
            return fail(OCTMPS_E_PARSE, "error reading ambient medium refractive index");
        octmps_region regs[OCTMPS_MAX_REGIONS];
        std::memset(regs, 0, sizeof(regs));
        regs[0].n = dtemp[0];
        for (int r = 1; r < nreg; ++r) {
            if (!readdoubles(4, dtemp, f.fp)) return fail(OCTMPS_E_PARSE, "error reading region data");
            regs[r].n = dtemp[0];
            regs[r].mua = dtemp[1];
            regs[r].mus = dtemp[2];
            regs[r].g = dtemp[3];
        }
        if (!readdoubles(3, dtemp, f.fp))
            return fail(OCTMPS_E_PARSE, "error reading probe position data parameters");
        if (i == run_index) {
            p->number_of_photons = nph;
            p->n_regions = (uint32_t)nreg;
            std::memcpy(p->regions, regs, sizeof(regs));
            p->probe_start_x = dtemp[0];
            p->probe_step_x = dtemp[1];
            p->probe_end_x = dtemp[2];
        }
    }
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_read_mci(const char *path, octmps_params *p) {
    // Bias_Read, src/bias.cpp:9-110: every real value goes through a float.
    if (!path || !p) return fail(OCTMPS_E_INVALID, "null argument");
    File f(path, "r");
    if (!f.fp) return fail(OCTMPS_E_IO, std::string("cannot open ") + path);
    float ft[5];
    int it[5];
    if (!readints(1, it, f.fp)) return fail(OCTMPS_E_PARSE, "error reading type of simulation");
    p->type_simulation = it[0];
    if (!readints(1, it, f.fp)) return fail(OCTMPS_E_PARSE, "error reading type of bias");
    p->type_bias = it[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading backward bias coefficient");
    p->backward_bias_coefficient = ft[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading minimum target depth");
    p->target_depth_min = ft[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading maximum target depth");
    p->target_depth_max = ft[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading coherence length source");
    p->coherence_length_source = ft[0];
    if (!readints(1, it, f.fp)) return fail(OCTMPS_E_PARSE, "error reading recording type");
    p->record_photons_flag = it[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading maximum collecting radius");
    p->max_collecting_radius = ft[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading maximum collecting angle");
    p->max_collecting_angle_deg = ft[0];
    if (!readfloats(1, ft, f.fp)) return fail(OCTMPS_E_PARSE, "error reading probability additional bias");
    p->probability_additional_bias = ft[0];
    if (!readfloats(1, ft, f.fp))
        return fail(OCTMPS_E_PARSE, "error reading maximum relative contribution per bin");
    p->max_relative_contribution = ft[0];
    const float const_max_fixed_depth = 0.12f;  // src/bias.cpp:92
    p->max_filtered_optical_depth =
        p->target_depth_max < const_max_fixed_depth ? (double)const_max_fixed_depth : p->target_depth_max;
    if (!(p->coherence_length_source > 0)) return fail(OCTMPS_E_RANGE, "coherence length must be > 0");
    // (short) truncation of NUM_SUBSTEPS_RESOLUTION * MFOD / Lc, src/bias.cpp:99-101.
    double nb = OCTMPS_NUM_SUBSTEPS_RESOLUTION * p->max_filtered_optical_depth / p->coherence_length_source;
    if (!(nb < 2147483647.0)) return fail(OCTMPS_E_RANGE, "bin count overflow");
    short bins = (short)(int)nb;
    if (bins <= 0) return fail(OCTMPS_E_RANGE, "bin count wraps the reference's short counter");
    p->num_depth_bins = bins;
    p->optical_depth_shift = fmod(p->target_depth_min, p->coherence_length_source);
    p->rnd_step_size_in_tissue = 0.0;
    return OCTMPS_OK;
}

extern "C" uint64_t octmps_tally_words(int32_t b) { return b > 0 ? tally_words(b) : 0; }

namespace {

// Exact value of one digit accumulator (include/octmps.h): value = sum_k w[k*B] * 2^(32k + exp0).
// Carries are propagated into 32-bit limbs, so the integer is held exactly; it is turned into
// a double only at the end (FP64 final normalisation).
struct BigNum {
    static constexpr int kLimbs = 24;  // 18 digits of up to 64 bits need at most 20 limbs
    uint32_t limb[kLimbs];
    int exp0;  // binary exponent of limb 0's unit

    BigNum(const uint64_t *w, int digits, uint64_t stride, int e0) : exp0(e0) {
        std::memset(limb, 0, sizeof limb);
        unsigned __int128 carry = 0;
        for (int k = 0; k < kLimbs; ++k) {
            unsigned __int128 t = carry + (k < digits ? (unsigned __int128)w[(uint64_t)k * stride] : 0);
            limb[k] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    // Subtracts m << shift (m < 2^64) from the integer; returns false (and leaves 0) if the
    // integer is smaller.
    bool sub(uint64_t m, int shift) {
        uint32_t d[kLimbs] = {};
        const int k = shift >> 5, r = shift & 31;
        const unsigned __int128 v = (unsigned __int128)m << r;
        for (int j = 0; j < 4 && k + j < kLimbs; ++j) d[k + j] = (uint32_t)(v >> (32 * j));
        int64_t borrow = 0;
        for (int j = 0; j < kLimbs; ++j) {
            int64_t t = (int64_t)limb[j] - (int64_t)d[j] - borrow;
            borrow = t < 0;
            limb[j] = (uint32_t)(t + (borrow << 32));
        }
        if (borrow) std::memset(limb, 0, sizeof limb);
        return !borrow;
    }
    double value() const {
        int h = kLimbs - 1;
        while (h >= 0 && limb[h] == 0) --h;
        if (h < 0) return 0.0;
        long double v = 0;
        const int lo = h >= 3 ? h - 3 : 0;
        for (int k = h; k >= lo; --k) v = v * 4294967296.0L + (long double)limb[k];
        return (double)ldexpl(v, 32 * lo + exp0);
    }
};

// FP32 fields of a contribution stored as binary64 (Max): c = m * 2^(e-150).
void fp32_fields(double c, uint64_t &m, int &e) {
    uint32_t b;
    float f = (float)c;
    std::memcpy(&b, &f, 4);
    e = (int)((b >> 23) & 0xffu);
    m = b & 0x7fffffu;
    if (e == 0) e = 1;
    else m |= 0x800000u;
}

struct Bin {
    double sum, sumsq, max;
    double fsum, fsumsq;  // Sum - Max and SumSq - Max^2, exact before rounding
    uint64_t count;
};

Bin get_bin(const uint64_t *t, int32_t B, int c, int b) {
    const uint64_t *base = t + 4 + (uint64_t)c * OCTMPS_TALLY_CLASS_WORDS * B;
    BigNum s(base + b, OCTMPS_TALLY_SUM_DIGITS, (uint64_t)B, OCTMPS_TALLY_SUM_EXP);
    BigNum q(base + (uint64_t)OCTMPS_TALLY_SUM_DIGITS * B + b, OCTMPS_TALLY_SQ_DIGITS, (uint64_t)B,
             OCTMPS_TALLY_SQ_EXP);
    Bin r;
    r.sum = s.value();
    r.sumsq = q.value();
    r.count = base[(uint64_t)(OCTMPS_TALLY_CLASS_WORDS - 2) * B + b];
    std::memcpy(&r.max, &base[(uint64_t)(OCTMPS_TALLY_CLASS_WORDS - 1) * B + b], sizeof(double));
    uint64_t m;
    int e;
    fp32_fields(r.max, m, e);
    if (r.max > 0) {
        s.sub(m, e + 10);
        q.sub(m * m, 2 * e + 20);
    }
    r.fsum = r.max > 0 ? s.value() : r.sum;
    r.fsumsq = r.max > 0 ? q.value() : r.sumsq;
    return r;
}

}  // namespace

extern "C" octmps_status octmps_tally_finalize(const uint64_t *tally, int32_t B, uint64_t n,
                                               double mrc, double *out) {
    // WriteReflectanceClassI_and_ClassII[_Filtered], src/octmps_io.cpp:302-420.
    if (!tally || !out || B <= 0) return fail(OCTMPS_E_INVALID, "bad finalize arguments");
    double N = (double)n;
    for (int c = 0; c < 2; ++c) {
        for (int b = 0; b < B; ++b) {
            Bin x = get_bin(tally, B, c, b);
            for (int f = 0; f < 2; ++f) {
                double sum = x.sum, sq = x.sumsq;
                if (f == 1 && x.max > mrc * sum) {
                    sum = x.fsum;
                    sq = x.fsumsq;
                }
                double mean = sum / N;
                double var = (sq / N - mean * mean) / (double)(n - 1);
                out[((f * 2 + c) * 2 + 0) * B + b] = mean;
                out[((f * 2 + c) * 2 + 1) * B + b] = sqrt(var);
            }
        }
    }
    return OCTMPS_OK;
}

extern "C" octmps_status octmps_write_ascan_outputs(const char *dir, const octmps_params *p,
                                                    double probe_x, const uint64_t *tally) {
    // WriteResult + Write_Bias_Parameters, src/octmps_io.cpp:237-437.
    if (!p || !tally) return fail(OCTMPS_E_INVALID, "null argument");
    std::string d = (dir && dir[0]) ? std::string(dir) + "/" : std::string();
    int32_t B = p->num_depth_bins;
    unsigned int nph = (unsigned int)p->number_of_photons;
    {
        File fp((d + "FilteredPhotons.mco").c_str(), "w");
        File bp((d + "BiasParam.mco").c_str(), "w");
        if (!bp.fp) return fail(OCTMPS_E_IO, "cannot open BiasParam.mco");
        FILE *o = bp.fp;
        fprintf(o, "%llu\t\t\t\t\t# Number of filtered photons\n", (unsigned long long)tally[0]);
        fprintf(o, "%llu\t\t\t\t\t# Number of Class I filtered photons\n", (unsigned long long)tally[1]);
        fprintf(o, "%llu\t\t\t\t\t# Number of Class II filtered photons\n", (unsigned long long)tally[2]);
        if (p->record_photons_flag) {
            fprintf(o, "\n\n\n\n");
            fprintf(o, "Description of fields in file FilteredPhotons.mco\n\n");
            fprintf(o, "================================================================\n");
            fprintf(o, "Weight | LikelihoodRatio | Radius | AngleRad | OpticalPath | MaxDepth");
        }
        fprintf(o, "%hd\t\t\t\t\t# Type of simulation\n", (short)p->type_simulation);
        fprintf(o, "%hd\t\t\t\t\t# Type of bias\n", (short)p->type_bias);
        fprintf(o, "%u\t\t\t\t\t# Number of photon packets\n", nph);
        fprintf(o, "%G\t\t\t\t\t# Backward bias coefficient\n", p->backward_bias_coefficient);
        fprintf(o, "%G\t\t\t\t\t# Target depth min [cm]\n", p->target_depth_min);
        fprintf(o, "%G\t\t\t\t\t# Target depth max [cm]\n", p->target_depth_max);
        fprintf(o, "%G\t\t\t\t\t# Coherence length source [cm]\n", p->coherence_length_source);
        fprintf(o, "%G\t\t\t\t\t# Max collecting radius [cm]\n", p->max_collecting_radius);
        fprintf(o, "%G\t\t\t\t\t# Max collecting angle [grad]\n", p->max_collecting_angle_deg);
        fprintf(o, "%hd\t\t\t\t\t# Record photons flag\n", (short)p->record_photons_flag);
        fprintf(o, "%G\t\t\t\t\t# Probability of additional bias\n", p->probability_additional_bias);
        fprintf(o, "%G\t\t\t\t\t# Maximum relative contribution to bin per post-processed sample\n",
                p->max_relative_contribution);
    }
    const char *names[2][2] = {{"ClassI_Scatt.out", "ClassII_Scatt.out"},
                               {"ClassI_ScattFilt.out", "ClassII_ScattFilt.out"}};
    for (int f = 0; f < 2; ++f) {
        File c1((d + names[f][0]).c_str(), "a");
        File c2((d + names[f][1]).c_str(), "a");
        if (!c1.fp || !c2.fp) return fail(OCTMPS_E_IO, "cannot open output CSV");
        for (long b = 0; b < B; ++b) {
            double depth = b * p->coherence_length_source / OCTMPS_NUM_SUBSTEPS_RESOLUTION +
                           p->optical_depth_shift;
            for (int c = 0; c < 2; ++c) {
                Bin x = get_bin(tally, B, c, (int)b);
                double sum = x.sum, sq = x.sumsq;
                if (f == 1 && x.max > p->max_relative_contribution * sum) {
                    sum = x.fsum;
                    sq = x.fsumsq;
                }
                double mean = sum / nph;
                double var = (sq / nph - mean * mean) / (nph - 1);
                fprintf(c == 0 ? c1.fp : c2.fp, "%G,%G,%G,%G,%llu\n", probe_x, depth, mean, sqrt(var),
                        (unsigned long long)x.count);
            }
        }
    }
    return OCTMPS_OK;
}
