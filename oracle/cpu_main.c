/*
 * cpu_main.c -- the reference's `cpu` program restated on the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle's command-line face): used by the
 * test suite to compare whole-program stdout with the GPU programs, and by the
 * CPU-baseline leg of the benchmark.  Follows src/cpu.cpp:274-352.
 *
 *   cpu [-f FILE] [-b BOUND] [-p] [--sort=bubble|stable] [--binary] [--time]
 *
 * Without -f the built-in 32-value sample of src/cpu.cpp:292-295 is indexed.
 * Values are read like `ifstream >> uint32_t` (src/cpu.cpp:298-302): skip
 * white space, optional sign, decimal digits; the first failed extraction
 * (no digits, overflow) ends the input.  --binary reads raw little-endian
 * uint32 instead (an extension).  --sort=bubble runs the literal O(n^2) sort
 * of src/cpu.cpp:180-197; the default sorts (value << 32 | rid) words, which
 * gives the same order.  --time prints "<us> us" for the pipeline to stderr.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "wah_oracle.h"

static uint32_t* read_text(const char* path, size_t* n_out) {
  FILE* f = fopen(path, "rb");
  size_t cap = 1 << 16, n = 0;
  uint32_t* v = (uint32_t*)malloc(cap * sizeof(uint32_t));
  *n_out = 0;
  if (!f) return v; /* an unreadable ifstream extracts nothing */
  int c = fgetc(f);
  for (;;) {
    while (c != EOF && isspace(c)) c = fgetc(f);
    if (c == EOF) break;
    int neg = 0;
    if (c == '+' || c == '-') { neg = (c == '-'); c = fgetc(f); }
    if (c == EOF || !isdigit(c)) break;
    uint64_t acc = 0;
    int overflow = 0;
    while (c != EOF && isdigit(c)) {
      acc = acc * 10u + (uint64_t)(c - '0');
      if (acc > 0xFFFFFFFFull) { overflow = 1; acc = 0xFFFFFFFFull; }
      c = fgetc(f);
    }
    if (overflow) break;
    if (n == cap) { cap *= 2; v = (uint32_t*)realloc(v, cap * sizeof(uint32_t)); }
    v[n++] = neg ? (uint32_t)(0u - (uint32_t)acc) : (uint32_t)acc;
  }
  fclose(f);
  *n_out = n;
  return v;
}

static uint32_t* read_binary(const char* path, size_t* n_out) {
  FILE* f = fopen(path, "rb");
  *n_out = 0;
  if (!f) return (uint32_t*)malloc(4);
  fseek(f, 0, SEEK_END);
  long bytes = ftell(f);
  fseek(f, 0, SEEK_SET);
  size_t n = (size_t)(bytes > 0 ? bytes : 0) / 4;
  uint32_t* v = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
  *n_out = fread(v, sizeof(uint32_t), n, f);
  fclose(f);
  return v;
}

int main(int argc, char** argv) {
  const char* file = NULL;
  uint32_t bound = 0;
  int print = 0, binary = 0, timing = 0, mode = WAH_ORACLE_SORT_PACKED;
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    if ((!strcmp(a, "-f") || !strcmp(a, "--data-file")) && i + 1 < argc) file = argv[++i];
    else if (!strncmp(a, "--data-file=", 12)) file = a + 12;
    else if ((!strcmp(a, "-b") || !strcmp(a, "--bound")) && i + 1 < argc) bound = (uint32_t)strtoul(argv[++i], 0, 10);
    else if (!strncmp(a, "--bound=", 8)) bound = (uint32_t)strtoul(a + 8, 0, 10);
    else if (!strcmp(a, "-p") || !strcmp(a, "--print")) print = 1;
    else if (!strcmp(a, "--binary")) binary = 1;
    else if (!strcmp(a, "--time")) timing = 1;
    else if (!strcmp(a, "--sort=bubble")) mode = WAH_ORACLE_SORT_BUBBLE;
    else if (!strcmp(a, "--sort=stable")) mode = WAH_ORACLE_SORT_PACKED;
    else { fprintf(stderr, "cpu: unknown argument '%s'\n", a); return 2; }
  }
  size_t n = 0;
  uint32_t* values;
  if (!file) {
    static const uint32_t sample[32] = {10, 7, 22, 6, 7, 1, 9, 42, 2, 5,
                                        13, 3, 2, 1, 0, 1, 18, 18, 3, 13,
                                        5, 9, 0, 3, 2, 19, 5, 23, 22, 10,
                                        6, 22};
    n = 32;
    values = (uint32_t*)malloc(sizeof(sample));
    memcpy(values, sample, sizeof(sample));
  } else {
    printf("Reading data from '%s' ... ", file);
    fflush(stdout);
    values = binary ? read_binary(file, &n) : read_text(file, &n);
  }
  printf("'%zu' values.\n", n);
  if (bound == 0 && n > 0) {
    for (size_t i = 0; i < n; ++i) if (values[i] > bound) bound = values[i];
  }
  printf("Maximum value is '%u'.\n", bound);

  size_t cap = n ? n : 1;
  uint32_t* keys = (uint32_t*)malloc(cap * 4);
  uint32_t* offsets = (uint32_t*)malloc(cap * 4);
  uint32_t* lengths = (uint32_t*)malloc(cap * 4);
  uint32_t* words = (uint32_t*)malloc(2 * cap * 4);
  wah_oracle_counts cnt;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (wah_oracle_build(values, n, 0, mode, keys, offsets, lengths, words, &cnt) != 0) {
    fprintf(stderr, "cpu: out of memory\n");
    return 1;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  printf("Merged to %llu values\n", (unsigned long long)cnt.k);
  printf("Index has %llu elements with %llu keys\n",
         (unsigned long long)cnt.length, (unsigned long long)cnt.keycnt);
  if (timing) {
    long long us = (long long)(t1.tv_sec - t0.tv_sec) * 1000000LL +
                   (t1.tv_nsec - t0.tv_nsec) / 1000;
    fprintf(stderr, "%lld us\n", us);
  }
  if (print) {
    uint32_t next = 0;
    size_t bufcap = 1 << 20;
    char* buf = (char*)malloc(bufcap);
    for (size_t i = 0; i < cnt.keycnt; ++i) {
      uint32_t key = keys[i];
      size_t need;
      while ((need = wah_oracle_decoded_bitmap(words + offsets[i], lengths[i], buf, bufcap)) > bufcap) {
        bufcap = need;
        buf = (char*)realloc(buf, bufcap);
      }
      if (key != next)
        for (uint32_t j = next; j < key; ++j) printf("%u\t\n", j);
      printf("%u\t%.*s\n", key, (int)need, buf);
      next = key + 1;
    }
    free(buf);
  }
  free(values); free(keys); free(offsets); free(lengths); free(words);
  return 0;
}
