// cc_state_driver.cpp -- runs the reference CPU solver
// (hydro-cal-fortan/calculate.cpp, compiled unmodified beside this file with
// its main renamed to hydro_cc_main) and writes its final state, which the
// solver itself never outputs.
//
// TEST INFRASTRUCTURE (the cross-check of the oracle against the reference
// CPU solver on the branches where the two discretizations agree).
//
// Usage (in a directory holding ./origin/changjiangkou/):
//   hydro_cc_state OUT.bin
// OUT.bin: int32 CELL, int32 steps, then double H, U, V, Z [CELL] of the
// state after the last step.  The solver keeps two time levels and writes
// step kt into level kt % 2 (calculate.cpp:1121-1133, read_data.hh:11-12).
#undef main
#include <cstdint>
#include <cstdio>
#include <vector>

using Vec2 = std::vector<std::vector<double>>;

// globals of calculate.cpp
extern Vec2 H, U, V, Z;
extern int CEL, MDT, DT, NDAYS;
int hydro_cc_main();

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s OUT.bin\n", argv[0]);
        return 2;
    }
    const int rc = hydro_cc_main();
    if (rc != 0) return rc;
    const int steps = MDT / DT;   // NDAYS = 1: kt = 1 .. MDT / DT
    const int lev = steps % 2;
    std::FILE *f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    const int32_t hdr[2] = {CEL, steps};
    std::fwrite(hdr, sizeof(int32_t), 2, f);
    for (const Vec2 *a : {&H, &U, &V, &Z}) std::fwrite((*a)[lev].data(), sizeof(double), CEL, f);
    std::fclose(f);
    return 0;
}
