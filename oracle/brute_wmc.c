/* ORACLE -- test infrastructure only.
 *
 * Brute-force weighted model count: the sum over all 2^n assignments of the product of
 * literal weights, over the assignments satisfying every clause.  This fixes the value
 * every correct contraction of the CNF's tensor network must produce
 * (tensor_network_constructions.py:7-41 builds that network; its full contraction is
 * this sum by construction).  Clauses are bit masks over <= 40 variables: a clause is
 * satisfied by x iff (x & pos) != 0 or (~x & neg) != 0.  The sum runs in long double.
 */
#include <stdint.h>

double brute_wmc(int n, int m, const uint64_t *pos, const uint64_t *neg,
                 const double *w_neg, const double *w_pos) {
    if (n < 0 || n > 40) return -1.0;
    long double total = 0.0L;
    uint64_t count = (uint64_t)1 << n;
    for (uint64_t x = 0; x < count; ++x) {
        int ok = 1;
        for (int c = 0; c < m; ++c) {
            if (!((x & pos[c]) || (~x & neg[c]))) { ok = 0; break; }
        }
        if (!ok) continue;
        long double w = 1.0L;
        for (int v = 0; v < n; ++v) w *= ((x >> v) & 1) ? w_pos[v] : w_neg[v];
        total += w;
    }
    return (double)total;
}
