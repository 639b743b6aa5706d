// ladder.cpp - the learned tempering ladder's training protocol, on the host,
// over the public C-ABI (ether_run, ether_pt_state, the moments, the MLP
// kernels of mlp.hip and ether_ladder_place).  Both drivers call it.
//
// No reference counterpart: the reference's ladder is linear in T
// (src/ether.f90:79-84) and its adaptive ladder is a commented stub
// (src/startup.f90:115-122).
//
// Which C.  The swap acceptance of neighbours (T1, T2) with Gaussian energy
// distributions is erfc(d/2), d = int sqrt(C)/T dT over the interval, so a
// measured acceptance a gives the length density lambda = 2 erfcinv(a)/|dT|
// (sqrt(C)/T at the interval; an interval with no accepted swap bounds it
// from below).  Far from equilibrium the fluctuation estimator of C is
// biased by drift, while this C is the one that sets the acceptance.
//
// Protocol.  `windows` windows of train_mcs/windows MCS; the first on the
// ladder the slots hold.  After each window the length map takes the
// intervals it measured, and (learning) the ladder is re-placed from the
// map, by the MLP surrogate (placement 0: trained on <e> and log C over the
// map's grid by the device kernels) or straight from the map (placement 1,
// the method without the network), and the slots are retuned in place.
// Then three validation windows of validate_mcs MCS, plain / learned /
// plain, each starting from the state the last left: the learned ladder is
// kept only if its round-trip cost Lambda is below the plain ladder's by more
// than twice the standard error of the difference (binomial spread of every
// pair's acceptance) and its lowest pair acceptance above the plain ladder's
// (plain: the mean of its two windows, which cancels a drift linear in time);
// otherwise the slots go back to the plain ladder.  A validation too short to
// tell the ladders apart so keeps the plain one.  learn = 0 runs the same windows on the plain ladder only,
// so both runs reach production after the same MCS.  Every swap passes the
// exact acceptance test, so the sampled distributions are unchanged.
#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "../../include/ether_gpu.h"

namespace ether {
int ctx_fail(ether_ctx* c, int code, const std::string& msg);
int ctx_R(ether_ctx* c);
int ctx_groups(const ether_ctx* c);
int ctx_snapshot(ether_ctx* c, void** out);
int ctx_restore(ether_ctx* c, void* snap);
void ctx_snapshot_free(void* snap);
}  // namespace ether

using namespace ether;

namespace {

// erfc(x) = y for y in (0, 2): bisection on the monotone erfc, to double precision
double erfcinv_d(double y) {
    double lo = -27.0, hi = 27.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (std::erfc(mid) > y) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

// Lambda = sum_k r_k / (1 - r_k), r the rejection rate of pair k (a pair that
// accepted nothing counts as half an acceptance).  Under the even/odd swap
// schedule a label needs about 2 + 2 Lambda swap rounds per round trip
// (Syed, Bouchard-Cote, Deligiannidis, Doucet, J. R. Stat. Soc. B 84 (2022)).
double round_trip_cost(const std::vector<double>& acc, const std::vector<double>& att) {
    double s = 0.0;
    for (size_t k = 0; k < acc.size(); ++k) {
        const double a = std::min(1.0, std::max(acc[k], 0.5 / std::max(att[k], 1.0)));
        s += (1.0 - a) / a;
    }
    return s;
}

// Variance of that estimate from the binomial spread of each pair's
// acceptance, A accepted of n attempts (delta method, d/da (1-a)/a = -1/a^2):
// sum_k (1 - a_k) / (a_k^3 n_k), with a = (A + 1)/(n + 2) (Laplace's rule)
// so that a pair with no accepted swap gets a finite, large variance.
double round_trip_var(const std::vector<double>& acc, const std::vector<double>& att) {
    double s = 0.0;
    for (size_t k = 0; k < acc.size(); ++k) {
        const double n = std::max(att[k], 1.0);
        const double a = (acc[k] * att[k] + 1.0) / (att[k] + 2.0);
        s += (1.0 - a) / (a * a * a * n);
    }
    return s;
}

struct LengthMap {
    std::vector<double> T, lam;
    // grid uniform in log T: lambda ~ sqrt(C)/T changes fastest at the cold end
    LengthMap(double lo, double hi, int n) : T(n), lam(n, std::numeric_limits<double>::quiet_NaN()) {
        for (int g = 0; g < n; ++g) T[g] = lo * std::pow(hi / lo, (double)g / (n - 1));
    }
    // a measured interval replaces what earlier windows said on its span; one
    // with no accepted swap only raises the value there to its lower bound
    void add(const std::vector<double>& temps, const std::vector<double>& acc,
             const std::vector<double>& att) {
        for (size_t k = 0; k + 1 < temps.size(); ++k) {
            const double at = std::max(att[k], 1.0);
            const bool bound = acc[k] * at < 0.5;
            const double a = std::min(0.999, std::max(acc[k], 0.5 / at));
            const double l = 2.0 * erfcinv_d(a) / std::fabs(temps[k + 1] - temps[k]);
            const double lo = std::min(temps[k], temps[k + 1]), hi = std::max(temps[k], temps[k + 1]);
            for (size_t g = 0; g < T.size(); ++g) {
                if (T[g] < lo || T[g] > hi) continue;
                if (bound) lam[g] = std::isnan(lam[g]) ? l : std::max(lam[g], l);
                else lam[g] = l;
            }
        }
    }
    // grid points no interval has covered take the nearest measured value
    std::vector<double> filled() const {
        std::vector<double> out(lam);
        for (size_t g = 0; g < out.size(); ++g) {
            if (!std::isnan(out[g])) continue;
            for (size_t d = 1; d < out.size(); ++d) {
                if (g >= d && !std::isnan(lam[g - d])) { out[g] = lam[g - d]; break; }
                if (g + d < out.size() && !std::isnan(lam[g + d])) { out[g] = lam[g + d]; break; }
            }
        }
        return out;
    }
};

// linear interpolation of (xs, ys), xs ascending, clamped at the ends
double interp(const std::vector<double>& xs, const std::vector<double>& ys, double x) {
    if (x <= xs.front()) return ys.front();
    if (x >= xs.back()) return ys.back();
    size_t k = std::upper_bound(xs.begin(), xs.end(), x) - xs.begin();
    const double w = (x - xs[k - 1]) / (xs[k] - xs[k - 1]);
    return ys[k - 1] + w * (ys[k] - ys[k - 1]);
}

// n temperatures on [lo, hi], hottest first, at equal steps of int lambda dT
// with lambda from the map (linear between grid points), on a fine grid
// uniform in log T like ether_ladder_place
void place_from_map(const LengthMap& m, double lo, double hi, int n, double* out) {
    const int G = 4096;
    const std::vector<double> lam = m.filled();
    std::vector<double> tg(G), cum(G, 0.0);
    for (int g = 0; g < G; ++g) tg[g] = lo * std::pow(hi / lo, (double)g / (G - 1));
    for (int g = 1; g < G; ++g)
        cum[g] = cum[g - 1] + 0.5 * (interp(m.T, lam, tg[g]) + interp(m.T, lam, tg[g - 1])) * (tg[g] - tg[g - 1]);
    const double total = cum[G - 1];
    for (int k = 0; k < n; ++k) {
        if (k == 0) { out[k] = hi; continue; }
        if (k == n - 1) { out[k] = lo; continue; }
        const double target = total * (double)(n - 1 - k) / (n - 1);
        int g = (int)(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
        g = std::max(1, std::min(G - 1, g));
        const double w = (target - cum[g - 1]) / std::max(cum[g] - cum[g - 1], 1e-300);
        out[k] = tg[g - 1] + w * (tg[g] - tg[g - 1]);
    }
}

struct Window {
    std::vector<double> acc, att, e;
};

}  // namespace

// One pass of the protocol (the windows, the validation, the ladder left for
// production); ether_ladder_train below runs it, and again on the plain
// ladder from the same start where the learned ladder is not kept.
static int train_pass(ether_ctx* c, int32_t n_temps, const double* temps, double kb,
                      int32_t first_step, int32_t train_mcs, int32_t windows,
                      int32_t validate_mcs, int32_t coa, int32_t ovrr_interval,
                      double ovrr_passes, int32_t pt_interval, int32_t learn,
                      int32_t placement, uint64_t mlp_seed, int32_t* swap_count,
                      double* T_out, double* report) {
    const int R = ctx_R(c);
    if (n_temps < 2 || !temps || !T_out || !swap_count || !(kb > 0.0) || R % n_temps != 0 ||
        train_mcs < 0 || windows < 1 || validate_mcs < 0 || pt_interval <= 0 ||
        (placement != 0 && placement != 1))
        return ctx_fail(c, ETHER_ERR_ARG, "ether_ladder_train: bad arguments");
    for (int k = 0; k + 1 < n_temps; ++k)
        if (!(temps[k] > temps[k + 1]) || !(temps[k + 1] > 0.0))
            return ctx_fail(c, ETHER_ERR_ARG, "ether_ladder_train: temperatures must fall, hottest first");
    const int n = n_temps, L = R / n;
    const double t_lo = temps[n - 1], t_hi = temps[0];
    std::vector<int32_t> gstart(L + 1);
    for (int g = 0; g <= L; ++g) gstart[g] = g * n;
    std::vector<double> plain(temps, temps + n), cur = plain, beta(R);
    int rc = ETHER_OK;
    int32_t step = first_step, sc = *swap_count;
    auto set_ladder = [&](const std::vector<double>& T) {
        for (int g = 0; g < L; ++g)
            for (int k = 0; k < n; ++k) beta[g * n + k] = 1.0 / (kb * T[k]);
        return ether_retune_betas(c, beta.data());
    };
    // one window: fresh swap counters and moments, every sample counted
    auto window = [&](int32_t len, Window& w) -> int {
        int r = ether_set_ladders(c, L, gstart.data());
        if (r) return r;
        if ((r = ether_reset_moments(c))) return r;
        if ((r = ether_run(c, step, len, step - 1, coa, ovrr_interval, ovrr_passes, pt_interval, &sc)))
            return r;
        step += len;
        std::vector<double> mom((size_t)R * ETHER_MOM_COUNT);
        std::vector<int64_t> acc(R), att(R);
        if ((r = ether_read_moments(c, mom.data()))) return r;
        if ((r = ether_pt_state(c, nullptr, acc.data(), att.data(), nullptr))) return r;
        w.acc.assign(n - 1, 0.0);
        w.att.assign(n - 1, 0.0);
        w.e.assign(n, 0.0);
        for (int k = 0; k < n; ++k)
            for (int g = 0; g < L; ++g) {
                const double* m = &mom[(size_t)(g * n + k) * ETHER_MOM_COUNT];
                w.e[k] += m[ETHER_MOM_E] / std::max(m[ETHER_MOM_N], 1.0) / L;
                if (k + 1 < n) {
                    w.acc[k] += (double)acc[g * n + k];
                    w.att[k] += (double)att[g * n + k];
                }
            }
        for (int k = 0; k + 1 < n; ++k) w.acc[k] = w.att[k] > 0.0 ? w.acc[k] / w.att[k] : 0.0;
        return ETHER_OK;
    };
    auto min_acc = [](const Window& w) { return *std::min_element(w.acc.begin(), w.acc.end()); };

    LengthMap map(t_lo, t_hi, 64);
    double loss = std::numeric_limits<double>::quiet_NaN();
    const int32_t win = std::max(train_mcs / windows, 1);
    for (int k = 0; k < windows && rc == ETHER_OK; ++k) {
        const int32_t len = k < windows - 1 ? win : train_mcs - win * (windows - 1);
        if (len <= 0) break;
        Window w;
        if ((rc = window(len, w))) break;
        map.add(cur, w.acc, w.att);
        if (!learn) continue;
        // <e> of the window on the map's grid, and the thermodynamic length
        // s(T) = int_lo^T lambda dT' of the map (trapezoid on its grid): each
        // measured interval fixes the length across it, 2 erfcinv(a), so s is
        // what the swaps measure; the network smooths it between the grid
        // points and the ladder takes equal steps of the smoothed s
        std::vector<double> order_T(cur.rbegin(), cur.rend()), order_e(w.e.rbegin(), w.e.rend());
        const std::vector<double> lam = map.filled();
        const int G = (int)map.T.size();
        std::vector<double> e(G), slen(G, 0.0);
        for (int g = 0; g < G; ++g) {
            e[g] = interp(order_T, order_e, map.T[g]);
            if (g) slen[g] = slen[g - 1] + 0.5 * (lam[g] + lam[g - 1]) * (map.T[g] - map.T[g - 1]);
        }
        std::vector<double> next(n);
        if (placement == 1) {
            place_from_map(map, t_lo, t_hi, n, next.data());
        } else {
            auto mean_sd = [](const std::vector<double>& v, double& mu, double& sd) {
                mu = 0.0;
                for (double x : v) mu += x;
                mu /= v.size();
                sd = 0.0;
                for (double x : v) sd += (x - mu) * (x - mu);
                sd = std::sqrt(sd / v.size());
            };
            double me, se;
            mean_sd(e, me, se);
            const double total = std::max(slen[G - 1], 1e-300);
            // the network's input is ln T, scaled to [-1, 1]; its outputs <e>
            // (standardised) and s / s_total scaled to [-1, 1]
            const double mid = 0.5 * (std::log(t_lo) + std::log(t_hi)), half = 0.5 * std::log(t_hi / t_lo);
            std::vector<float> x(G), y(2 * G), wt(2 * G, 1.0f);
            for (int g = 0; g < G; ++g) {
                x[g] = (float)((std::log(map.T[g]) - mid) / half);
                y[2 * g] = (float)((e[g] - me) / (se + 1e-30));
                y[2 * g + 1] = (float)(2.0 * slen[g] / total - 1.0);
            }
            float fl = 0.f;
            if ((rc = ether_mlp_init(c, mlp_seed))) break;
            if ((rc = ether_mlp_train(c, G, x.data(), y.data(), wt.data(), 12000, 3e-3f, &fl))) break;
            loss = fl;
            if ((rc = ether_ladder_place(c, mid, half, 1.0, t_lo, t_hi, n, ETHER_PLACE_LOG_T | ETHER_PLACE_LENGTH,
                                         next.data())))
                break;
        }
        cur = next;
        if ((rc = set_ladder(cur))) break;
    }
    // validation: plain, learned, plain
    double cost_p = 0.0, cost_l = 0.0, min_p = 0.0, min_l = 0.0, sd = 0.0;
    int kept = 0;
    if (rc == ETHER_OK && validate_mcs > 0) {
        const std::vector<double> cand = cur;
        Window a, b, d;
        if (!(rc = set_ladder(plain)) && !(rc = window(validate_mcs, a)) &&
            !(rc = set_ladder(learn ? cand : plain)) && !(rc = window(validate_mcs, b)) &&
            !(rc = set_ladder(plain)) && !(rc = window(validate_mcs, d))) {
            cost_p = 0.5 * (round_trip_cost(a.acc, a.att) + round_trip_cost(d.acc, d.att));
            min_p = 0.5 * (min_acc(a) + min_acc(d));
            cost_l = round_trip_cost(b.acc, b.att);
            min_l = min_acc(b);
            sd = std::sqrt(0.25 * (round_trip_var(a.acc, a.att) + round_trip_var(d.acc, d.att)) +
                           round_trip_var(b.acc, b.att));
            kept = learn && cost_p - cost_l > 2.0 * sd && min_l > min_p;
            cur = kept ? cand : plain;
            if (kept) rc = set_ladder(cur);
        }
    } else if (rc == ETHER_OK) {
        kept = learn ? 1 : 0;
    }
    if (rc == ETHER_OK) rc = ether_set_ladders(c, L, gstart.data());
    if (rc != ETHER_OK) return rc;
    for (int k = 0; k < n; ++k) T_out[k] = cur[k];
    *swap_count = sc;
    if (report) {
        report[0] = cost_p;
        report[1] = cost_l;
        report[2] = min_p;
        report[3] = min_l;
        report[4] = kept;
        report[5] = loss;
        report[6] = (double)(step - first_step);
        report[7] = sd;
    }
    return ETHER_OK;
}

// Where validation does not keep the learned ladder, the context goes back to
// its state at entry (ctx_snapshot) and runs the same windows on the plain
// ladder: production then continues exactly as a run with learn = 0 would
// (every draw is a function of the sweep and swap counters), so a rejected
// learned ladder costs the repeated windows and changes no result.  The
// report keeps the validation figures of the learned attempt.
extern "C" int ether_ladder_train(ether_ctx* c, int32_t n_temps, const double* temps, double kb,
                                  int32_t first_step, int32_t train_mcs, int32_t windows,
                                  int32_t validate_mcs, int32_t coa, int32_t ovrr_interval,
                                  double ovrr_passes, int32_t pt_interval, int32_t learn,
                                  int32_t placement, uint64_t mlp_seed, int32_t* swap_count,
                                  double* T_out, double* report) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!swap_count) return ctx_fail(c, ETHER_ERR_ARG, "ether_ladder_train: bad arguments");
    if (!learn || validate_mcs <= 0)
        return train_pass(c, n_temps, temps, kb, first_step, train_mcs, windows, validate_mcs, coa,
                          ovrr_interval, ovrr_passes, pt_interval, learn, placement, mlp_seed, swap_count,
                          T_out, report);
    void* snap = nullptr;
    int rc = ctx_snapshot(c, &snap);
    if (rc) return rc;
    const int32_t sc0 = *swap_count;
    double rep[8];
    rc = train_pass(c, n_temps, temps, kb, first_step, train_mcs, windows, validate_mcs, coa, ovrr_interval,
                    ovrr_passes, pt_interval, 1, placement, mlp_seed, swap_count, T_out, rep);
    if (rc == ETHER_OK && rep[4] == 0.0) {
        double again[8];
        *swap_count = sc0;
        rc = ctx_restore(c, snap);
        if (rc == ETHER_OK)
            rc = train_pass(c, n_temps, temps, kb, first_step, train_mcs, windows, validate_mcs, coa,
                            ovrr_interval, ovrr_passes, pt_interval, 0, placement, mlp_seed, swap_count,
                            T_out, again);
    }
    ctx_snapshot_free(snap);
    if (rc == ETHER_OK && report)
        for (int k = 0; k < 8; ++k) report[k] = rep[k];
    return rc;
}
