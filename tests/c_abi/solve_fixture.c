/* A plain C client of the two C ABIs (include/qiotk.h, include/qmc.h):
 *
 *   solve_fixture TARGET PARAMS.json INPUT.json
 *
 * checks the device RNG against the Philox4x32-10 known answer for a zero
 * counter and key through qmc.h, then solves the input file with the params
 * document through qiotk.h (Runner::configure + get_run_output,
 * app/runner.cc:215-260, 513-688) and prints the response JSON on stdout.
 * Exit status: the solve's status code (0 = ok); 3 when the known answer
 * differs. */
#include <stdio.h>
#include <stdlib.h>

#include "qiotk.h"
#include "qmc.h"

static char* slurp(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char* s = (char*)malloc((size_t)n + 1);
  if (s && fread(s, 1, (size_t)n, f) != (size_t)n) {
    free(s);
    s = NULL;
  }
  if (s) s[n] = 0;
  fclose(f);
  return s;
}

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s TARGET PARAMS.json INPUT.json\n", argv[0]);
    return 2;
  }
  /* qmc.h: Philox4x32-10 of counter 0, key 0 (Random123 known answer) */
  qmc_ctx ctx = NULL;
  if (qmc_ctx_create(0, &ctx) != QMC_OK) {
    fprintf(stderr, "qmc_ctx_create: %s\n", qmc_last_error());
    return 2;
  }
  const uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
  uint32_t words[4] = {0, 0, 0, 0};
  const int prc = qmc_philox(ctx, 1, ctr, key, words);
  qmc_ctx_destroy(ctx);
  if (prc != QMC_OK || words[0] != 0x6627E8D5u || words[1] != 0xE169C58Du || words[2] != 0xBC57AC4Cu ||
      words[3] != 0x9B00DBD8u) {
    fprintf(stderr, "philox known answer differs: %08x %08x %08x %08x\n", words[0], words[1], words[2], words[3]);
    return 3;
  }
  /* qiotk.h: one solve of the input file */
  char* params = slurp(argv[2]);
  if (!params) {
    fprintf(stderr, "cannot read %s\n", argv[2]);
    return 2;
  }
  qiotk_engine e = NULL;
  if (qiotk_engine_create(0, &e) != 0) {
    fprintf(stderr, "qiotk_engine_create: %s\n", qiotk_last_error());
    free(params);
    return 2;
  }
  char* out = NULL;
  const int rc = qiotk_solve_file(e, argv[1], params, argv[3], &out);
  printf("%s\n", out ? out : "");
  qiotk_free(out);
  qiotk_engine_destroy(e);
  free(params);
  return rc;
}
