// hexo_cli.cpp -- the `hexo` command line for the pricing path.
//
// Keeps the reference's grammar (src/Main.cpp:58-96, src/inc/Main.h:12):
//   hexo -p asian all <underlying> [<volume type> <# volume>]   prices every call of every chain
//   hexo -c <underlying> [<volume type> <# volume>]             calibrates the Heston parameters
//   hexo -d <underlying> [<volume type> <# volume>]             lists the market data (no GPU)
//   hexo -t rng | -t simulation                                  self-tests on the GPU
//   hexo -h
// The reference downloads quotes and chains over HTTPS and keeps calibrated parameters in
// sqlite (src/Main.cpp:64-86). This CLI reads them offline from $HEXO_DATA_DIR (default ./data):
// the Tradier JSON responses saved under <underlying>/ (market_data.hpp), or a compact chain file
// <underlying>.chains; parameters live in HParams.db there (the reference's table) or, without
// libsqlite3, in <underlying>.params. As in the reference, -p uses stored parameters (a `params`
// line of the chain file, else the store) and calibrates and stores them when there are none.
// Chain file, one record per line:
//   quote <bid>
//   params <v0> <theta> <rho> <kappa> <sigma>          (HParams order)
//   expiry <days to expiry>
//   call <strike> <bid> <ask> <volume> <open_interest>  (belongs to the last expiry)
// Chains with T <= 0.1 years are dropped and calls are kept when the selected volume field is
// >= <# volume> (src/WebAPI.cpp:62-84,101-105). The output line per option has the reference's
// columns (src/Main.cpp:89-93). Environment: HEXO_PATHS (1e5), HEXO_STEPS (1e3),
// HEXO_PRECISION (fp64 | fp32, default fp64), HEXO_SEED (1); variance reduction, off by default:
// HEXO_CONTROL_VARIATE (none | geo = regression: the regression coefficient | geo-unit:
// coefficient 1), HEXO_MARTINGALE (0 | 1).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hexo/HCalibration.h"
#include "../../include/hexo/HSimulation.h"
#include "market_data.hpp"

namespace {

const char kHelp[] =
    "hexo [{-c <underlying> [<volume type> <# volume>]}...] [{-d <underlying> [<volume type> <# volume>]}...] "
    "[{-p asian all <underlying> [<volume type> <# volume>]}...] [{-t rng|simulation}...] [-h]";
const double kTradingDays = (double)(5 * (365 / 7) + 365 % 7);
const double kExpLowerBound = 0.1;
const double kYearlyRiskFree = 0.0005;

struct UnderlyingData {
    double S = 0;
    bool have_params = false;
    HParams p{};
    std::list<options_chain> chains;
};

std::string data_dir() {
    const char* dir = std::getenv("HEXO_DATA_DIR");
    return dir ? dir : "data";
}

UnderlyingData read_chains(const std::string& sym, const std::string& vol_type, long vol_n) {
    const std::string path = data_dir() + "/" + sym + ".chains";
    std::ifstream in(path);
    if (!in) {
        // the API's JSON responses saved under <underlying>/
        const char* today = std::getenv("HEXO_TODAY");
        hexo_cli::MarketData md = hexo_cli::read_tradier_dir(data_dir() + "/" + sym, vol_type, vol_n, today ? today : "");
        UnderlyingData d;
        d.S = md.S;
        d.chains = std::move(md.chains);
        return d;
    }
    UnderlyingData d;
    std::string line;
    bool keep = false;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag) || tag[0] == '#') continue;
        if (tag == "quote") {
            ss >> d.S;
        } else if (tag == "params") {
            ss >> d.p.v_0 >> d.p.v_m >> d.p.rho >> d.p.kappa >> d.p.sigma;
            d.have_params = !ss.fail();
        } else if (tag == "expiry") {
            unsigned days = 0;
            ss >> days;
            const double T = days / kTradingDays;
            keep = T > kExpLowerBound;
            if (keep) d.chains.emplace_back(days, T);
        } else if (tag == "call") {
            double strike, bid, ask;
            long volume, oi;
            ss >> strike >> bid >> ask >> volume >> oi;
            if (ss.fail()) throw std::runtime_error("malformed call line: " + line);
            if (!keep) continue;
            const long v = vol_type == "volume" ? volume : oi;
            if (v < vol_n) continue;
            option o{};
            o.price = ask;
            o.bid = bid;
            o.strike = strike;
            o.volume = v;
            options_chain& c = d.chains.back();
            c.options.push_back(o);
            c.min_strike = std::min(c.min_strike, strike);
            c.max_strike = std::max(c.max_strike, strike);
        }
    }
    if (!(d.S > 0)) throw std::runtime_error(path + ": missing quote");
    return d;
}

// calibrate() with the reference's progress lines (src/Main.cpp:69-72, src/HCalibration.cpp:81,109).
HParams calibrate_and_store(const std::string& sym, const UnderlyingData& d) {
    unsigned n = 0;
    for (const auto& c : d.chains) n += (unsigned)c.options.size();
    std::cout << "setup completed\t# observations: " << n << '\n';
    const HParams p = calibrate(d.S, d.chains);
    std::cout << "v_0: " << p.v_0 << "\tv_m: " << p.v_m << "\trho: " << p.rho << "\tkappa: " << p.kappa
              << "\tsigma: " << p.sigma << '\n';
    hexo_cli::store_params(data_dir(), sym, p);
    return p;
}

double norm_cdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }

double bs_call(double S, double K, double vol, double T) {
    const double d1 = (std::log(S / K) + (kYearlyRiskFree + 0.5 * vol * vol) * T) / (vol * std::sqrt(T));
    return S * norm_cdf(d1) - K * std::exp(-kYearlyRiskFree * T) * norm_cdf(d1 - vol * std::sqrt(T));
}

// Implied volatility by bracketing and bisection to 1e-6; -1 below the arbitrage bound, -2 above S
// (the printed column of the reference, src/BSM.cpp:22-50).
double imp_vol(double S, const option& o, double T) {
    if (o.price <= S - std::exp(-kYearlyRiskFree * T) * o.strike) return -1.0;
    if (o.price >= S) return -2.0;
    double lo = 1.0 / 64, hi = lo;
    if (bs_call(S, o.strike, lo, T) < o.price) {
        while (bs_call(S, o.strike, hi, T) < o.price) hi *= 2;
        lo = hi / 2;
    } else {
        while (bs_call(S, o.strike, lo, T) >= o.price) lo *= 0.5;
        hi = lo * 2;
    }
    double mid = lo;
    while (hi - lo >= 1e-6) {
        mid = 0.5 * (lo + hi);
        (bs_call(S, o.strike, mid, T) < o.price ? lo : hi) = mid;
    }
    return mid;
}

// The DOWNLOAD listing of the reference (src/Main.cpp:97-103).
void list_options(const UnderlyingData& d) {
    for (const auto& c : d.chains)
        for (const option& o : c.options)
            std::cout << "S: " << std::setw(10) << std::right << std::setfill(' ') << std::fixed << std::setprecision(2)
                      << d.S << "\tstrike: " << o.strike << "\tbid: " << o.bid << "\task: " << o.price
                      << "\tvolume: " << o.volume << "\timp vol: " << imp_vol(d.S, o, c.time_to_expiry)
                      << "\tlb: " << d.S - std::exp(-kYearlyRiskFree * c.time_to_expiry) * o.strike
                      << "\texpiry time: " << c.time_to_expiry * kTradingDays << '\n';
}

unsigned env_u(const char* name, unsigned dflt) {
    const char* v = std::getenv(name);
    return v ? (unsigned)std::strtod(v, nullptr) : dflt;
}

template <class Scheme>
std::vector<ffloat> run_price(const UnderlyingData& d, unsigned paths, unsigned n_opts, unsigned steps) {
    return HSimulation::price<Scheme>(d.p, d.S, d.chains, paths, n_opts, steps);
}

// The scheme from the environment: precision, Andersen's martingale correction, and the
// geometric-average control variate (the regression coefficient, or coefficient 1 on request).
template <typename acc, class Policy>
std::vector<ffloat> run_scheme(bool mart, const UnderlyingData& d, unsigned paths, unsigned n_opts, unsigned steps) {
    return mart ? run_price<HSimulation::HQEAndersonMartingale<acc, Policy>>(d, paths, n_opts, steps)
                : run_price<HSimulation::HQEAnderson<acc, Policy>>(d, paths, n_opts, steps);
}

template <typename acc>
std::vector<ffloat> run_policy(const std::string& cv, bool mart, const UnderlyingData& d, unsigned paths,
                               unsigned n_opts, unsigned steps) {
    if (cv == "geo" || cv == "regression") return run_scheme<acc, AAsianCallGeometricCV>(mart, d, paths, n_opts, steps);
    if (cv == "geo-unit") return run_scheme<acc, AAsianCallGeometricCVUnitCoefficient>(mart, d, paths, n_opts, steps);
    if (!cv.empty() && cv != "none")
        throw std::runtime_error("HEXO_CONTROL_VARIATE must be none, geo (= regression) or geo-unit");
    return run_scheme<acc, AAsianCallNonAdaptive>(mart, d, paths, n_opts, steps);
}

int price_all(const std::string& sym, const std::string& vol_type, long vol_n) {
    UnderlyingData d = read_chains(sym, vol_type, vol_n);
    if (!d.have_params && !hexo_cli::fetch_params(data_dir(), sym, d.p)) d.p = calibrate_and_store(sym, d);
    unsigned n_opts = 0;
    for (const auto& c : d.chains) n_opts += (unsigned)c.options.size();
    std::cout << "params, v0: " << d.p.v_0 << "\tv_m: " << d.p.v_m << "\trho: " << d.p.rho << "\tkappa: " << d.p.kappa
              << "\tsigma: " << d.p.sigma << std::endl;
    if (n_opts == 0) return 0;
    const unsigned paths = env_u("HEXO_PATHS", 100000), steps = env_u("HEXO_STEPS", 1000);
    HSimulation::set_seed(env_u("HEXO_SEED", 1));
    const char* prec = std::getenv("HEXO_PRECISION");
    const bool fp32 = prec && std::string(prec) == "fp32";
    const char* cv_env = std::getenv("HEXO_CONTROL_VARIATE");
    const std::string cv = cv_env ? cv_env : "";
    const bool mart = env_u("HEXO_MARTINGALE", 0) != 0;
    std::vector<ffloat> res = fp32 ? run_policy<float>(cv, mart, d, paths, n_opts, steps)
                                   : run_policy<ffloat>(cv, mart, d, paths, n_opts, steps);
    unsigned i = 0;
    for (const options_chain& c : d.chains)
        for (const option& o : c.options)
            std::cout << "S: " << std::setw(10) << std::right << std::setfill(' ') << std::fixed << std::setprecision(2)
                      << d.S << "\tstrike: " << o.strike << "\tbid: " << o.bid << "\task: " << o.price
                      << "\tasian-option-price: " << res[i++] << "\tvolume: " << o.volume
                      << "\timp vol: " << imp_vol(d.S, o, c.time_to_expiry)
                      << "\tlb: " << d.S - std::exp(-kYearlyRiskFree * c.time_to_expiry) * o.strike
                      << "\texpiry time: " << c.time_to_expiry * kTradingDays << '\n';
    return 0;
}

// Moments of 2^26 uniforms and normals drawn on the GPU (src/UnitTest.cpp:528-564 draws 2^28).
// The reference's rng_test (src/UnitTest.cpp:528-564) at its sample size: 2^28 uniforms and 2^28
// normals (the FP32 path's inverse normal) from the Philox stream, mean and variance of each against the
// exact moments. The reference's bound 1e-4 is kept where it is at least five standard errors at
// this size (the uniform's mean, 5.5 SE, and variance, 22 SE); the normal's mean and variance
// have standard errors 6.1e-5 and 8.6e-5, so they are held to five of those.
int test_rng() {
    hexo_ctx* ctx = HSimulation::detail::context();
    const uint64_t chunk_paths = 1 << 12, chunk_words = 1 << 12, n_chunks = 16;  // 16 x 2^24 = 2^28
    std::vector<uint32_t> w(chunk_paths * chunk_words);
    std::vector<double> z(w.size());
    double s = 0, s2 = 0, zs = 0, zs2 = 0;
    for (uint64_t c = 0; c < n_chunks; ++c) {
        if (hexo_rng_fill(ctx, 1, 0, c * chunk_paths, chunk_paths, 0, chunk_words, w.data()) != HEXO_OK ||
            hexo_normals(ctx, w.data(), w.size(), HEXO_FP32, z.data()) != HEXO_OK)
            return 1;
        double cs = 0, cs2 = 0, czs = 0, czs2 = 0;  // per-chunk sums, then added: short summation chains
        for (size_t i = 0; i < w.size(); ++i) {
            const double u = (w[i] + 0.5) * 0x1p-32;
            cs += u;
            cs2 += u * u;
            czs += z[i];
            czs2 += z[i] * z[i];
        }
        s += cs;
        s2 += cs2;
        zs += czs;
        zs2 += czs2;
    }
    const double n = (double)(n_chunks * w.size());
    const double mu = s / n, var = (s2 - n * mu * mu) / (n - 1);
    const double zm = zs / n, zv = (zs2 - n * zm * zm) / (n - 1);
    std::printf("sample size %.0f: uniform mean %.7f var %.7f; normal mean %.7f var %.7f\n", n, mu, var, zm, zv);
    const bool ok = std::fabs(mu - 0.5) < 1e-4 && std::fabs(var - 1.0 / 12) < 1e-4 &&
                    std::fabs(zm) < 5.0 / std::sqrt(n) && std::fabs(zv - 1.0) < 5.0 * std::sqrt(2.0 / n);
    std::printf("RNG test %s\n", ok ? "passed" : "FAILED");
    return ok ? 0 : 1;
}

// FP64 and FP32 kernels on the same draws agree; the Asian price lies in its no-arbitrage range.
int test_simulation() {
    std::list<options_chain> chains;
    chains.emplace_back(261, 1.0);
    option o{};
    o.strike = 100.0;
    chains.back().options.push_back(o);
    const HParams p{0.04, 0.04, -0.7, 2.0, 0.3};
    std::vector<ffloat> se64, se32;
    auto a64 = HSimulation::price_with_errors<HSimulation::HQEAnderson<double, AAsianCallNonAdaptive>>(
        p, 100.0, chains, 200000, 1, 252, &se64);
    auto a32 = HSimulation::price_with_errors<HSimulation::HQEAnderson<float, AAsianCallNonAdaptive>>(
        p, 100.0, chains, 200000, 1, 252, &se32);
    const double rel = std::fabs(a32[0] - a64[0]) / a64[0];
    std::printf("asian call FP64 %.6f +- %.6f, FP32 %.6f (rel diff %.2e)\n", a64[0], se64[0], a32[0], rel);
    const bool ok = rel < 1e-4 && a64[0] > 0 && a64[0] < 100.0;
    std::printf("Simulation test %s\n", ok ? "passed" : "FAILED");
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char* argv[]) {
    try {
        int cur = 1;
        if (argc == 1) {
            std::cout << kHelp << std::endl;
            return 0;
        }
        while (cur < argc) {
            const std::string a = argv[cur];
            if (a == "-p" || a == "--price") {
                if (cur + 3 >= argc || std::string(argv[cur + 1]) != "asian" || std::string(argv[cur + 2]) != "all")
                    throw std::runtime_error(kHelp);
                const std::string sym = argv[cur + 3];
                std::string vol_type = "open_interest";
                long vol_n = 0;
                cur += 4;
                if (cur + 1 < argc && argv[cur][0] != '-') {
                    vol_type = argv[cur];
                    vol_n = std::stol(argv[cur + 1]);
                    cur += 2;
                }
                if (price_all(sym, vol_type, vol_n)) return 1;
            } else if (a == "-c" || a == "--calibrate" || a == "-d" || a == "--download") {
                if (cur + 1 >= argc) throw std::runtime_error(kHelp);
                const std::string sym = argv[cur + 1];
                std::string vol_type = "open_interest";
                long vol_n = 0;
                cur += 2;
                if (cur + 1 < argc && argv[cur][0] != '-') {
                    vol_type = argv[cur];
                    vol_n = std::stol(argv[cur + 1]);
                    cur += 2;
                }
                const UnderlyingData d = read_chains(sym, vol_type, vol_n);
                if (a == "-d" || a == "--download") {
                    list_options(d);
                } else {
                    std::cout << "Options data downloaded and parsed\n";
                    calibrate_and_store(sym, d);
                }
            } else if (a == "-t" || a == "--test") {
                if (cur + 1 >= argc) throw std::runtime_error(kHelp);
                const std::string t = argv[cur + 1];
                int rc = t == "rng" ? test_rng() : t == "simulation" ? test_simulation() : -1;
                if (rc < 0) throw std::runtime_error(kHelp);
                if (rc) return rc;
                cur += 2;
            } else if (a == "-h" || a == "--help") {
                std::cout << kHelp << std::endl;
                ++cur;
            } else {
                throw std::runtime_error(kHelp);
            }
        }
    } catch (const std::exception& e) {
        std::cerr << "hexo: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
