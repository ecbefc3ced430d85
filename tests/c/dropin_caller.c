/*
 * dropin_caller.c -- a C program written against the reference's own header
 * layout (include/lmie_interface.h, src/lmie_interface.h:20-157) and linked
 * to liblmie_amd.so, as an existing LMie caller would be.  It solves one
 * size distribution through lmie_solution (src/lmie_interface.c:199-250) with
 * alloc_out = 1 and prints every scalar output and the expansion
 * coefficients with 17 significant digits.
 *
 *   dropin_caller dist n_int2 n_quad lambda mr mi a1 a2 r1 r2
 */
#include <stdio.h>
#include <stdlib.h>

#include "lmie_interface.h"

int main(int argc, char **argv) {
     if (argc != 11) {
          fprintf(stderr, "usage: %s dist n_int2 n_quad lambda mr mi a1 a2 r1 r2\n", argv[0]);
          return 2;
     }
     lmie_in_data in = {0};
     lmie_out_data out = {0};
     in.calc_gc = 1;
     in.dist_type = (enum size_dist_type) atoi(argv[1]);
     in.n_int2 = atoi(argv[2]);
     in.n_quad = atoi(argv[3]);
     in.lambda = atof(argv[4]);
     in.mr = atof(argv[5]);
     in.mi = atof(argv[6]);
     in.a1 = atof(argv[7]);
     in.a2 = atof(argv[8]);
     in.r1 = atof(argv[9]);
     in.r2 = atof(argv[10]);
     in.accuracy = -1.;
     if (lmie_solution(&in, &out, 1, 0, 1, 0)) {
          fprintf(stderr, "lmie_solution failed\n");
          return 1;
     }
     printf("n_coef %d\n", out.n_coef);
     printf("scalars %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e\n",
            out.r1, out.r2, out.norm, out.reff, out.veff, out.gavg, out.vavg, out.ravg, out.rvw,
            out.cext, out.csca, out.cbak, out.g);
     for (int k = 0; k < 6; ++k) {
          printf("gc%d", k);
          for (int l = 0; l < out.n_coef; ++l) printf(" %.17e", out.gc[k][l]);
          printf("\n");
     }
     lmie_out_free(&out, 0);
     return 0;
}
