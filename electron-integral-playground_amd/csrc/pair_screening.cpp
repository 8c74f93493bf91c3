// Native enumeration and pre-screening of primitive pairs (host code, multithreaded).
//
// Restates the loop of reference pair_utility.py:78-124 (form_primitive_pair_list):
//   * shell pairs (i, j) over all shells in order, admitted when l_i < l_j, or l_i == l_j and
//     ao_start_i <= ao_start_j (pair_utility.py:87-91);
//   * for every primitive pair, the distance is shortened CUMULATIVELY inside the primitive loops,
//     d -= sqrt(l_i / 2 / a) per i primitive and d -= sqrt(l_j / 2 / b) per j primitive, then clamped with
//     Python's max(d, 0.0) (pair_utility.py:85, 99, 107, 109) -- so a pair's distance depends on the
//     primitives visited before it within the shell pair;
//   * the bound 2 |c_a| g_a |c_b| g_b exp(-ab/(a+b) d^2) max((pi/p)^1.5, 2^0.25 (pi/p)^1.25) with
//     g = (l / (2 e a))^(l/2) (pair_utility.py:95-114).
//
// The distance arithmetic (subtraction, correctly rounded sqrt, the clamp) is IEEE-exact and identical to
// the reference's.  The bound uses libm exp/pow, which may differ by a few ulp from numpy's vectorised
// transcendental functions, so this pass is a conservative PRE-screen: it keeps every pair whose bound is
// at least threshold * (1 - margin) and returns, per candidate, the indices and the clamped distance.  The
// caller re-evaluates the exact bound of the candidates with the reference's numpy operations and applies
// the exact keep test and the stable ordering, which makes the final list the reference's bit for bit.
// The atom-pair distance |A - B| comes from the caller (numpy's norm), for the same reason.

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace {

struct ScreenInput {
    int n_shell;
    const int* l;
    const int* ao_start;
    const int* atom;
    const int* prim_off;     // [n_shell + 1]
    const double* exps;
    const double* coefs;
    const double* atom_dist; // [n_atom][n_atom]
    int n_atom;
    double cut;              // threshold * (1 - margin)
};

inline bool admitted(const ScreenInput& in, int i, int j)
{
    if (in.l[i] > in.l[j]) return false;
    if (in.l[i] == in.l[j] && in.ao_start[i] > in.ao_start[j]) return false;
    return true;
}

// Visit the candidates of shell row i in reference order; f(j, ip, jp, d) per candidate.
template <class F>
void visit_row(const ScreenInput& in, int i, F&& f)
{
    const double kPi = 3.141592653589793;
    const double kE = 2.718281828459045;
    const double kTwoQuarter = std::pow(2.0, 0.25);
    const int li = in.l[i];
    for (int j = 0; j < in.n_shell; ++j) {
        if (!admitted(in, i, j)) continue;
        const int lj = in.l[j];
        double d = in.atom_dist[(long long) in.atom[i] * in.n_atom + in.atom[j]];
        for (int ip = 0; ip < in.prim_off[i + 1] - in.prim_off[i]; ++ip) {
            const double a = in.exps[in.prim_off[i] + ip];
            double cA = std::fabs(in.coefs[in.prim_off[i] + ip]);
            if (li > 0) cA *= std::pow(li / (2.0 * kE * a), li / 2.0);
            d -= std::sqrt(li / 2.0 / a);
            for (int jp = 0; jp < in.prim_off[j + 1] - in.prim_off[j]; ++jp) {
                const double b = in.exps[in.prim_off[j] + jp];
                double cB = std::fabs(in.coefs[in.prim_off[j] + jp]);
                if (lj > 0) cB *= std::pow(lj / (2.0 * kE * b), lj / 2.0);
                d -= std::sqrt(lj / 2.0 / b);
                d = (0.0 > d) ? 0.0 : d;
                const double g = std::exp(-a * b / (a + b) * d * d);
                const double cab = 2.0 * cA * cB;
                const double ov = std::pow(kPi / (a + b), 1.5);
                const double sw = std::pow(kPi / (a + b), 1.25) * kTwoQuarter;
                const double ub = cab * g * std::max(ov, sw);
                if (ub >= in.cut) f(j, ip, jp, d);
            }
        }
    }
}

template <class F>
void parallel_rows(int n_rows, F&& f)
{
    int nt = (int) std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    nt = std::max(1, std::min(nt, n_rows));
    std::vector<std::thread> pool;
    // interleaved rows: shells of one angular class sit together and have similar cost
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] { for (int i = t; i < n_rows; i += nt) f(i); });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" long long eip_screen_primitive_pairs(
    int n_shell, const int* shell_l, const int* shell_ao_start, const int* shell_atom, const int* shell_prim_off,
    const double* prim_exp, const double* prim_coef, const double* atom_dist, int n_atom, double threshold,
    double margin, long long capacity, int* out_i_shell, int* out_j_shell, int* out_i_prim, int* out_j_prim,
    double* out_dist)
{
    if (n_shell <= 0 || !shell_l || !shell_ao_start || !shell_atom || !shell_prim_off || !prim_exp || !prim_coef ||
        !atom_dist || n_atom <= 0 || margin < 0.0 || margin >= 1.0)
        return -1;
    const ScreenInput in{n_shell, shell_l, shell_ao_start, shell_atom, shell_prim_off, prim_exp, prim_coef,
                         atom_dist, n_atom, threshold * (1.0 - margin)};
    std::vector<long long> count(n_shell + 1, 0);
    parallel_rows(n_shell, [&](int i) {
        long long c = 0;
        visit_row(in, i, [&](int, int, int, double) { ++c; });
        count[i + 1] = c;
    });
    for (int i = 0; i < n_shell; ++i) count[i + 1] += count[i];
    const long long total = count[n_shell];
    if (capacity < total || !out_i_shell || !out_j_shell || !out_i_prim || !out_j_prim || !out_dist)
        return total;
    parallel_rows(n_shell, [&](int i) {
        long long k = count[i];
        visit_row(in, i, [&](int j, int ip, int jp, double d) {
            out_i_shell[k] = i;
            out_j_shell[k] = j;
            out_i_prim[k] = ip;
            out_j_prim[k] = jp;
            out_dist[k] = d;
            ++k;
        });
    });
    return total;
}
