// A caller with the reference driver's shape (with_space_optimization/main.cu:
// argv passed through, rank 0, the two out-parameters printed), linked against
// the binding in symbfact_min_id.cpp.
#include <cstdio>

typedef unsigned long long ull_t;
void symbfact_min_id(int argc, char** argv, int myrank, ull_t& fill_count, double& time);

int main(int argc, char** argv) {
    if (argc > 4) printf("Number of GPUs used: %s\n", argv[4]);
    ull_t fills = 0;
    double ms = 0;
    symbfact_min_id(argc, argv, 0, fills, ms);
    printf("caller: fill_count=%llu time_ms=%.3f\n", fills, ms);
    return fills > 0 ? 0 : 3;
}
