// gsofa_cli.cpp -- drop-in command line with the reference's positional argv:
//   gsofa colbeg colend csr N_GPU MEM_GB [max_supernode_size [N_GPU_Node]]
// (with_space_optimization/symbfact.cuh:1557-1571; the supernode variant adds
// the two optional knobs, with_space_optimization/with_supernodes/
// symbfact.cuh:1562-1578).  MEM_GB is the GiB pool for labels, markers and
// queues; N_GPU virtual GPUs run their cyclic source shares one after the
// other on this device, as in the reference (README.md:56-57).  Prints the
// reference's summary lines (with_space_optimization/symbfact.cuh:2524-2530).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gsofa.h"

static int usage(const char* argv0) {
    fprintf(stderr,
            "usage: %s colbeg colend csr N_GPU MEM_GB [max_supernode_size [N_GPU_Node]]\n"
            "       [--chunk C] [--share G] [--counts-only] [--spill] [--order-nd LEAF] [--device D]\n",
            argv0);
    return 1;
}

int main(int argc, char** argv) {
    const char* pos[7] = {0};
    int npos = 0;
    gsofa_options opt;
    gsofa_options_default(&opt);
    int order_leaf = 0;  // > 0: reorder the input by nested dissection first
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--chunk") && i + 1 < argc) opt.chunk = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--share") && i + 1 < argc) opt.virtual_gpu_id = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) opt.device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--counts-only")) opt.flags |= GSOFA_FLAG_COUNTS_ONLY;
        else if (!strcmp(argv[i], "--spill")) opt.flags |= GSOFA_FLAG_SPILL;
        else if (!strcmp(argv[i], "--order-nd") && i + 1 < argc) order_leaf = atoi(argv[++i]);
        else if (npos < 7) pos[npos++] = argv[i];
        else return usage(argv[0]);
    }
    if (npos < 5) return usage(argv[0]);
    opt.n_virtual_gpu = atoi(pos[3]);
    double mem_gb = atof(pos[4]);
    if (opt.n_virtual_gpu < 1 || mem_gb < 0) return usage(argv[0]);
    opt.mem_budget_bytes = (int64_t)(mem_gb * 1024.0 * 1024.0 * 1024.0);
    if (npos >= 6) opt.max_supernode_size = atoi(pos[5]);
    printf("Number of GPUs used: %d\n", opt.n_virtual_gpu);
    gsofa_csr m;
    int rc = gsofa_read_ascii(pos[0], pos[1], pos[2], &m);
    if (rc) {
        fprintf(stderr, "gsofa: reading input failed: %s\n", gsofa_strerror(rc));
        return 2;
    }
    gsofa_matrix A = {m.n, m.nnz, m.colbeg, m.colend, m.csr};
    if (order_leaf > 0) {  // factor P A P^T instead of A
        std::vector<int32_t> perm((size_t)(m.n > 0 ? m.n : 1));
        gsofa_csr b;
        rc = gsofa_order_nd(&A, order_leaf, perm.data());
        if (!rc) rc = gsofa_permute(&A, perm.data(), &b);
        if (rc) {
            fprintf(stderr, "gsofa: ordering failed: %s\n", gsofa_strerror(rc));
            gsofa_csr_free(&m);
            return 2;
        }
        gsofa_csr_free(&m);
        m = b;
        A = {m.n, m.nnz, m.colbeg, m.colend, m.csr};
        printf("Reordered by nested dissection (leaf %d)\n", order_leaf);
    }
    gsofa_result r;
    rc = gsofa_symbfact(&A, &opt, &r);
    if (rc) {
        fprintf(stderr, "gsofa: %s\n", gsofa_strerror(rc));
        gsofa_csr_free(&m);
        return 3;
    }
    printf("vertex count: %d |Total_allocation: %g GB  |beginning_N_src_group: %d  |ending_N_src_group: %d"
           "  |N_virtual_GPUs: %d\n",
           m.n, mem_gb, r.chunk, r.chunk, opt.n_virtual_gpu);
    printf("Total Number of fill-ins detected: %llu\n", (unsigned long long)r.fill_count);
    printf("\tSlowest virtual GPU takes: %g ms\n", r.share_ms_max);
    printf("\tFastest virtual GPU takes: %g ms\n", r.share_ms_min);
    printf("TE_alg: %llu  |R|: %llu  edge checks: %llu  levels: %llu  overflow rescans: %llu\n",
           (unsigned long long)r.te_alg, (unsigned long long)r.reach, (unsigned long long)r.edge_checks,
           (unsigned long long)r.levels, (unsigned long long)r.rescans);
    if (r.n_super >= 0) printf("Supernodes: %d\n", r.n_super);
    printf("Device time: %g ms (traversal %g ms)\n", r.ms, r.traverse_ms);
    gsofa_result_free(&r);
    gsofa_csr_free(&m);
    return 0;
}
