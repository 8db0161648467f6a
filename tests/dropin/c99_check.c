/* The C ABI from strict C99 (what a cgo, JNI or ctypes binding sees): host
 * calls, then one device call.  With three file arguments (the reference's
 * ASCII triple) the device call factors that matrix; otherwise a 3x3 one.
 * Prints: n_virtual_gpu virtual_gpu_id order_rc permute_rc symbfact_rc fill_count */
#include <stdio.h>
#include "gsofa.h"
int main(int argc, char** argv) {
    gsofa_options o;
    gsofa_options_default(&o);
    int32_t colbeg[3] = {0, 2, 5}, colend[3] = {2, 5, 7}, csr[7] = {0, 1, 0, 1, 2, 1, 2};
    gsofa_matrix A = {3, 7, colbeg, colend, csr};
    gsofa_csr F;
    int have_file = 0;
    if (argc == 4) {
        int rr = gsofa_read_ascii(argv[1], argv[2], argv[3], &F);
        if (rr != GSOFA_OK) {
            fprintf(stderr, "read: %s\n", gsofa_strerror(rr));
            return 2;
        }
        A.n = F.n;
        A.nnz = F.nnz;
        A.colbeg = F.colbeg;
        A.colend = F.colend;
        A.csr = F.csr;
        have_file = 1;
    }
    int32_t perm[3];
    int32_t tiny[3] = {0, 2, 5}, tend[3] = {2, 5, 7};
    gsofa_matrix T = {3, 7, tiny, tend, csr};
    int rc = gsofa_order_nd(&T, 64, perm);
    gsofa_csr B;
    int rp = gsofa_permute(&T, perm, &B);
    if (rp == GSOFA_OK) gsofa_csr_free(&B);
    gsofa_result r;
    unsigned long long fills = 0;
    int rs = gsofa_symbfact(&A, &o, &r);
    if (rs == GSOFA_OK) {
        fills = (unsigned long long)r.fill_count;
        gsofa_result_free(&r);
    }
    printf("%d %d %d %d %d %llu\n", o.n_virtual_gpu, o.virtual_gpu_id, rc, rp, rs, fills);
    if (have_file) gsofa_csr_free(&F);
    return 0;
}
