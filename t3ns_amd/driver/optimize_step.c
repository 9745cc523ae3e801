/*
 * t3ns_optimize_step: one optimisation step through the C ABI, in the call
 * sequence of the reference's optimize_siteTensor
 * (src/optimize_network.c:188-226):
 *
 *     init_Heffdata(&mv_dat, operators, &msiteObj);      t3ns_heff_init
 *     diagonal = make_diagonal(&mv_dat);                 t3ns_make_diagonal
 *     sparse_eigensolve(msiteObj.blocks.tel, &energy,    t3ns_sparse_eigensolve
 *         size, 30, 2, davidson_rtl, davidson_max_its,     with t3ns_matvecT3NS
 *         diagonal, matvecT3NS, &mv_dat, "D");
 *     destroy_Heffdata(&mv_dat);                         t3ns_heff_destroy
 *     safe_free(diagonal);                               free
 *
 * then the same solve with the search space and the matvec on the device
 * (t3ns_davidson_dev with t3ns_heff_dev_matvec).  The step's inputs come
 * from the host layer (include/t3ns_host.h): a network file, a model
 * (NN_HUBBARD or an FCIDUMP), symmetries, target, bond dimension and the
 * step's sites, with seeded operators.
 *
 * Usage:
 *   t3ns_optimize_step --netw FILE (--hubbard T,U | --fcidump FILE)
 *       --sym U1,SU2 --ts 10,0 --D 24 --sites 0,1,10,2 [--seed 7]
 *       [--guess FILE] [--tol 1e-10] [--its 200] [--out FILE]
 * --guess: the initial vector as raw float64 (default: all ones);
 * --out: the eigenvector of the reference-shaped solve as raw float64.
 * Prints one JSON line: size, energies, statuses of both solves.
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "t3ns_amd.h"
#include "t3ns_host.h"

static char* read_file(const char* path, size_t* len)
{
        FILE* f = fopen(path, "rb");
        if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        char* buf = malloc((size_t) n + 1);
        if (!buf || fread(buf, 1, (size_t) n, f) != (size_t) n) {
                fprintf(stderr, "cannot read %s\n", path);
                exit(2);
        }
        buf[n] = '\0';
        fclose(f);
        if (len) *len = (size_t) n;
        return buf;
}

static int parse_ints(const char* s, int* out, int cap)
{
        int n = 0;
        while (*s && n < cap) {
                char* end;
                out[n++] = (int) strtol(s, &end, 10);
                s = *end == ',' ? end + 1 : end;
                if (end == s && *s) break;
        }
        return n;
}

static int symmetry_id(const char* name)
{
        /* enum symmetrygroup order (include/symmetries.h:34) */
        static const char* names[] = {"Z2", "U1", "SU2", "C1", "Ci", "C2", "Cs", "D2", "C2v",
                                      "C2h", "D2h"};
        for (int i = 0; i < (int) (sizeof names / sizeof names[0]); ++i)
                if (!strcmp(name, names[i])) return i;
        fprintf(stderr, "unknown symmetry %s\n", name);
        exit(2);
}

static void check(int rc, const char* what)
{
        if (rc != T3NS_OK) {
                fprintf(stderr, "%s failed (%d): %s\n", what, rc, t3ns_last_error());
                exit(1);
        }
}

int main(int argc, char** argv)
{
        const char *netw = NULL, *fcidump = NULL, *guess = NULL, *out = NULL;
        double hub[2] = {1.0, 4.0};
        int hubbard = 0, sgs[T3NS_MAX_SYMMETRIES], ts[T3NS_MAX_SYMMETRIES], sites[4];
        int nsym = 0, nts = 0, nsites = 0, D = 0, its = 200;
        uint64_t seed = 7;
        double tol = 1e-10;
        for (int i = 1; i + 1 < argc; i += 2) {
                const char *k = argv[i], *v = argv[i + 1];
                if (!strcmp(k, "--netw")) netw = v;
                else if (!strcmp(k, "--fcidump")) fcidump = v;
                else if (!strcmp(k, "--hubbard")) {
                        hubbard = 1;
                        if (sscanf(v, "%lf,%lf", &hub[0], &hub[1]) != 2) { fprintf(stderr, "--hubbard T,U\n"); return 2; }
                } else if (!strcmp(k, "--sym")) {
                        char buf[256];
                        snprintf(buf, sizeof buf, "%s", v);
                        for (char* t = strtok(buf, ","); t && nsym < T3NS_MAX_SYMMETRIES; t = strtok(NULL, ","))
                                sgs[nsym++] = symmetry_id(t);
                } else if (!strcmp(k, "--ts")) nts = parse_ints(v, ts, T3NS_MAX_SYMMETRIES);
                else if (!strcmp(k, "--D")) D = atoi(v);
                else if (!strcmp(k, "--sites")) nsites = parse_ints(v, sites, 4);
                else if (!strcmp(k, "--seed")) seed = strtoull(v, NULL, 10);
                else if (!strcmp(k, "--guess")) guess = v;
                else if (!strcmp(k, "--tol")) tol = atof(v);
                else if (!strcmp(k, "--its")) its = atoi(v);
                else if (!strcmp(k, "--out")) out = v;
                else { fprintf(stderr, "unknown option %s\n", k); return 2; }
        }
        if (!netw || (!fcidump && !hubbard) || nsym == 0 || nts != nsym || D <= 0 || nsites < 2) {
                fprintf(stderr, "usage: t3ns_optimize_step --netw FILE (--hubbard T,U | --fcidump FILE) "
                        "--sym U1,SU2 --ts 10,0 --D 24 --sites 0,1,10,2 [--seed 7] [--guess FILE] "
                        "[--tol 1e-10] [--its 200] [--out FILE]\n");
                return 2;
        }
        if (t3ns_device_count() < 1) { fprintf(stderr, "no HIP device\n"); return 3; }

        /* the solve context and the step (the reference's globals and
         * o_dat, src/optimize_network.c:138-177) */
        char* netw_text = read_file(netw, NULL);
        t3ns_ctx* ctx = NULL;
        if (fcidump) {
                char* fci = read_file(fcidump, NULL);
                check(t3ns_ctx_create_qc(&ctx, netw_text, fci, sgs, ts, nsym, D, 2), "t3ns_ctx_create_qc");
                free(fci);
        } else {
                check(t3ns_ctx_create(&ctx, netw_text, "NN_HUBBARD", hub, 2, sgs, ts, nsym, D, 2),
                      "t3ns_ctx_create");
        }
        free(netw_text);
        t3ns_step* step = NULL;
        check(t3ns_step_random(&step, ctx, sites, nsites, seed), "t3ns_step_random");
        t3ns_heff_desc desc;
        t3ns_rops_desc ops[3];
        t3ns_site_desc site;
        check(t3ns_step_heff_desc(step, &desc, ops, &site), "t3ns_step_heff_desc");
        const int size = (int) t3ns_step_size(step);

        double* tel = malloc(sizeof(double) * (size_t) size);   /* msiteObj.blocks.tel */
        double* x0 = malloc(sizeof(double) * (size_t) size);
        if (!tel || !x0) { fprintf(stderr, "out of memory\n"); return 1; }
        if (guess) {
                size_t len;
                char* g = read_file(guess, &len);
                if (len != sizeof(double) * (size_t) size) {
                        fprintf(stderr, "guess has %zu bytes, the step needs %d doubles\n", len, size);
                        return 2;
                }
                memcpy(x0, g, len);
                free(g);
        } else {
                for (int i = 0; i < size; ++i) x0[i] = 1.0;
        }
        memcpy(tel, x0, sizeof(double) * (size_t) size);

        /* optimize_siteTensor, src/optimize_network.c:196-226 */
        t3ns_heff* mv_dat = NULL;
        check(t3ns_heff_init(&mv_dat, &desc, ops, &site, T3NS_FP64), "init_Heffdata");
        double* diagonal = t3ns_make_diagonal(mv_dat);
        if (!diagonal) check(T3NS_ERR_INVALID, "make_diagonal");
        double energy = 0;
        const int status = t3ns_sparse_eigensolve(tel, &energy, size, 30, 2, tol, its, diagonal,
                                                  t3ns_matvecT3NS, mv_dat, "D");

        /* the same solve on the device: t3ns_davidson_dev with the Heff as
         * its device callback */
        double *dx = NULL, *dd = NULL;
        if (hipMalloc((void**) &dx, sizeof(double) * (size_t) size) != hipSuccess ||
            hipMalloc((void**) &dd, sizeof(double) * (size_t) size) != hipSuccess) {
                fprintf(stderr, "hipMalloc failed\n");
                return 1;
        }
        hipMemcpy(dx, x0, sizeof(double) * (size_t) size, hipMemcpyHostToDevice);
        hipMemcpy(dd, diagonal, sizeof(double) * (size_t) size, hipMemcpyHostToDevice);
        double energy_dev = 0;
        int its_dev = 0;
        const int status_dev = t3ns_davidson_dev(size, dx, dd, t3ns_heff_dev_matvec, mv_dat, NULL,
                                                 &energy_dev, 30, 2, tol, its, &its_dev);
        hipFree(dx);
        hipFree(dd);

        t3ns_heff_destroy(mv_dat);                              /* destroy_Heffdata */
        free(diagonal);                                         /* safe_free(diagonal) */

        if (out) {
                FILE* f = fopen(out, "wb");
                if (!f || fwrite(tel, sizeof(double), (size_t) size, f) != (size_t) size) {
                        fprintf(stderr, "cannot write %s\n", out);
                        return 1;
                }
                fclose(f);
        }
        printf("{\"size\": %d, \"energy\": %.17g, \"status\": %d, \"energy_dev\": %.17g, "
               "\"status_dev\": %d, \"its_dev\": %d}\n",
               size, energy, status, energy_dev, status_dev, its_dev);
        free(tel);
        free(x0);
        t3ns_step_destroy(step);
        t3ns_ctx_destroy(ctx);
        return status < 0 || status_dev < 0 ? 1 : 0;
}
