// The reference-side binding: the reference's driver entry point
//     void symbfact_min_id(int argc, char** argv, int myrank, ull_t& fill_count, double& time)
// (with_space_optimization/symbfact.cuh:1553, called from with_space_optimization/main.cu:31)
// re-implemented on the C ABI of include/gsofa.h.  A maintainer links this in
// place of the reference's body and keeps the caller unchanged.
#include <cstdio>
#include <cstdlib>
#include "gsofa.h"

typedef unsigned long long ull_t;

void symbfact_min_id(int argc, char** argv, int myrank, ull_t& fill_count, double& time) {
    // argv: colbeg colend csr N_GPU MEM_GB [max_supernode_size [N_GPU_Node]]
    // (with_space_optimization/symbfact.cuh:1557-1571; with_supernodes/symbfact.cuh:1562-1578)
    fill_count = 0;
    time = 0;
    if (argc < 6) { fprintf(stderr, "gsofa: colbeg colend csr N_GPU MEM_GB expected\n"); return; }
    gsofa_csr A;
    int rc = gsofa_read_ascii(argv[1], argv[2], argv[3], &A);
    if (rc) { fprintf(stderr, "gsofa: %s\n", gsofa_strerror(rc)); return; }
    gsofa_options opt;
    gsofa_options_default(&opt);
    opt.n_virtual_gpu = atoi(argv[4]);
    opt.mem_budget_bytes = (int64_t)(atof(argv[5]) * (1ll << 30));
    if (argc > 6) opt.max_supernode_size = atoi(argv[6]);
    opt.device = myrank;                 // one process per GPU
    gsofa_matrix M = {A.n, A.nnz, A.colbeg, A.colend, A.csr};
    gsofa_result r;
    rc = gsofa_symbfact(&M, &opt, &r);
    if (rc == GSOFA_OK) {
        printf("Total Number of fill-ins detected: %llu\n", (ull_t)r.fill_count);
        fill_count = r.fill_count;       // the total over all shares (the reference's
                                         //  out-parameter held only the last share)
        time = r.ms;                     // milliseconds, as the reference's
                                         //  (with_space_optimization/symbfact.cuh:2420)
        gsofa_result_free(&r);
    } else {
        fprintf(stderr, "gsofa: %s\n", gsofa_strerror(rc));
    }
    gsofa_csr_free(&A);
}
