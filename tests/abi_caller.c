/* A compiled C caller of include/reaxamd.h: the calls a LAMMPS adapter makes for one force evaluation, in C99 against the
 * plain C ABI (no C++ and no Python in between).
 *
 *   abi_caller FFIELD ELEMENTS CASE [PRECISION]
 *
 * ELEMENTS is a comma list (one element per LAMMPS type); CASE is the case text of tests/golden/make_golden.py (origin, lattice
 * vectors a b c in LAMMPS form, periodic flags, "atoms N" + "tag type x y z" lines, "charges" q...). It runs rxa_create ->
 * rxa_pair_style -> rxa_pair_coeff -> rxa_set_box -> rxa_set_atoms -> rxa_set_charges -> rxa_compute and prints the 13 energy
 * terms, the 14 pvector slots, the virial and the gradient with 17 significant digits. Exit status: 0 success, 3 no device
 * (rxa_create returned RXA_EDEVICE), 1 any other failure (the context's error text on stderr). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "reaxamd.h"

static char* read_file(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char* buf = (char*)malloc((size_t)n + 1);
  if (buf && fread(buf, 1, (size_t)n, f) != (size_t)n) { free(buf); buf = NULL; }
  if (buf) buf[n] = 0;
  fclose(f);
  return buf;
}

static int fail(rxa_context* ctx, const char* what, int rc) {
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, ctx ? rxa_last_error(ctx) : "");
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: abi_caller FFIELD ELEMENTS CASE [PRECISION]\n");
    return 2;
  }
  const char* precision = argc > 4 ? argv[4] : "fp64";
  char* ffield = read_file(argv[1]);
  if (!ffield) { fprintf(stderr, "cannot read %s\n", argv[1]); return 1; }
  const char* elements[16];
  int ntypes = 0;
  char* el = (char*)malloc(strlen(argv[2]) + 1);
  if (!el) return 1;
  strcpy(el, argv[2]);
  for (char* tok = strtok(el, ","); tok && ntypes < 16; tok = strtok(NULL, ",")) elements[ntypes++] = tok;

  double origin[3] = {0, 0, 0}, vec[3][3] = {{0}};
  int periodic[3] = {1, 1, 1};
  long n = 0;
  int* types = NULL;
  int64_t* tags = NULL;
  double *x = NULL, *q = NULL;
  FILE* in = fopen(argv[3], "r");
  if (!in) { fprintf(stderr, "cannot read %s\n", argv[3]); return 1; }
  char word[64];
  while (fscanf(in, "%63s", word) == 1) {
    if (!strcmp(word, "origin")) {
      if (fscanf(in, "%lf %lf %lf", &origin[0], &origin[1], &origin[2]) != 3) return 1;
    } else if (!strcmp(word, "a") || !strcmp(word, "b") || !strcmp(word, "c")) {
      double* v = vec[word[0] - 'a'];
      if (fscanf(in, "%lf %lf %lf", &v[0], &v[1], &v[2]) != 3) return 1;
    } else if (!strcmp(word, "periodic")) {
      if (fscanf(in, "%d %d %d", &periodic[0], &periodic[1], &periodic[2]) != 3) return 1;
    } else if (!strcmp(word, "atoms")) {
      if (fscanf(in, "%ld", &n) != 1 || n <= 0) return 1;
      types = (int*)calloc((size_t)n, sizeof(int));
      tags = (int64_t*)calloc((size_t)n, sizeof(int64_t));
      x = (double*)calloc(3 * (size_t)n, sizeof(double));
      q = (double*)calloc((size_t)n, sizeof(double));
      for (long i = 0; i < n; ++i) {
        long long t;
        if (fscanf(in, "%lld %d %lf %lf %lf", &t, &types[i], &x[3 * i], &x[3 * i + 1], &x[3 * i + 2]) != 5) return 1;
        tags[i] = (int64_t)t;
      }
    } else if (!strcmp(word, "charges")) {
      for (long i = 0; i < n; ++i)
        if (fscanf(in, "%lf", &q[i]) != 1) return 1;
    }
  }
  fclose(in);
  if (!n) { fprintf(stderr, "no atoms in %s\n", argv[3]); return 1; }
  /* LAMMPS box: lo = origin, hi = origin + diagonal, tilt = (xy, xz, yz) */
  const double lo[3] = {origin[0], origin[1], origin[2]};
  const double hi[3] = {origin[0] + vec[0][0], origin[1] + vec[1][1], origin[2] + vec[2][2]};
  const double tilt[3] = {vec[1][0], vec[2][0], vec[2][1]};

  rxa_context* ctx = NULL;
  int rc = rxa_create(&ctx, 0);
  if (rc == RXA_EDEVICE) { fprintf(stderr, "no device\n"); return 3; }
  if (rc != RXA_OK) return fail(NULL, "rxa_create", rc);
  const char* style[3] = {"NULL", "precision", precision};
  if ((rc = rxa_pair_style(ctx, 3, style, NULL)) != RXA_OK) return fail(ctx, "rxa_pair_style", rc);
  int map[16];
  if ((rc = rxa_pair_coeff(ctx, ffield, ntypes, elements, map)) != RXA_OK) return fail(ctx, "rxa_pair_coeff", rc);
  if ((rc = rxa_set_box(ctx, lo, hi, tilt, periodic)) != RXA_OK) return fail(ctx, "rxa_set_box", rc);
  if ((rc = rxa_set_atoms(ctx, n, types, tags, x)) != RXA_OK) return fail(ctx, "rxa_set_atoms", rc);
  if ((rc = rxa_set_charges(ctx, q)) != RXA_OK) return fail(ctx, "rxa_set_charges", rc);
  double e[RXA_NTERMS], vir[6], pv[14];
  double* grad = (double*)calloc(3 * (size_t)n, sizeof(double));
  if ((rc = rxa_compute(ctx, RXA_COMPUTE_REBUILD, e, grad, vir)) != RXA_OK) return fail(ctx, "rxa_compute", rc);
  rxa_pvector(e, pv);
  for (int k = 0; k < RXA_NTERMS; ++k) printf("energy %d %.17g\n", k, e[k]);
  for (int k = 0; k < 14; ++k) printf("pvector %d %.17g\n", k, pv[k]);
  for (int k = 0; k < 6; ++k) printf("virial %d %.17g\n", k, vir[k]);
  for (long i = 0; i < n; ++i) printf("grad %ld %.17g %.17g %.17g\n", i, grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]);
  rxa_destroy(ctx);
  free(grad); free(types); free(tags); free(x); free(q); free(el); free(ffield);
  return 0;
}
