// cc_interp_driver.cpp -- evaluates BOUNDRYinterp of the reference CPU solver
// (hydro-cal-fortan/calculate.cpp:1107-1116, compiled unmodified beside this
// file with its main renamed) on tables read from stdin.
//
// TEST INFRASTRUCTURE (a check of the oracle's boundary interpolation
// against the reference's compiled code).
//
// stdin: n, then n pairs `time value`, then m, then m query times.
// stdout: m lines, each the interpolated value printed with %.17g.
// Queries must lie below the last table time: the reference's loop reads one
// entry past the table when a query reaches it.
#include <cstdio>
#include <vector>

double BOUNDRYinterp(double THOURS, int NZQSTEMP, std::vector<double> ZQSTIME, std::vector<double> ZQSTEMP);

int main() {
    int n = 0, m = 0;
    if (std::scanf("%d", &n) != 1 || n < 2) return 2;
    std::vector<double> t(n), v(n);
    for (int i = 0; i < n; i++)
        if (std::scanf("%lf %lf", &t[i], &v[i]) != 2) return 2;
    if (std::scanf("%d", &m) != 1) return 2;
    for (int q = 0; q < m; q++) {
        double x;
        if (std::scanf("%lf", &x) != 1) return 2;
        if (!(x < t[n - 1])) return 3;
        std::printf("%.17g\n", BOUNDRYinterp(x, n, t, v));
    }
    return 0;
}
