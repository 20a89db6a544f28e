/* TEST INFRASTRUCTURE ONLY (the oracle).  Never linked into the product.
 *
 * Raw COMMON-block dumper used by ref_harness.f.  The harness registers
 * each block's base address and end address (LOC() of its first and last
 * members); mhdump_ writes them all to one file:
 *
 *   "MHDUMP1\0"  int32 real_size  int32 nblocks  double loop_seconds
 *   per block:   char name[16]  int64 nbytes  bytes...
 *
 * With MH_NODUMP set in the environment only the header is written.
 * loop_seconds is the wall time from the MHATEXIT call (just before LOOP)
 * to the exit hook, i.e. the reference's own time-marching loop.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAXBLK 64
static struct { char name[16]; const void *base; int64_t nbytes; } blk[MAXBLK];
static int nblk = 0;
static int32_t real_size = 4;
static char exit_path[256];
static double t_loop0 = -1.0;

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void copy_name(char *dst, size_t cap, const char *src, long len) {
  size_t n = (size_t)len < cap - 1 ? (size_t)len : cap - 1;
  memcpy(dst, src, n);
  dst[n] = 0;
  while (n > 0 && dst[n - 1] == ' ') dst[--n] = 0;
}

void mhbegin_(const int32_t *rs) { real_size = *rs; nblk = 0; }

void mhblk_(const char *name, const int64_t *first, const int64_t *end, long len) {
  if (nblk >= MAXBLK) { fprintf(stderr, "ref_dump: too many blocks\n"); exit(3); }
  memset(blk[nblk].name, 0, 16);
  copy_name(blk[nblk].name, 16, name, len);
  blk[nblk].base = (const void *)(intptr_t)(*first);
  blk[nblk].nbytes = *end - *first;
  nblk++;
}

static void dump_to(const char *path, double secs) {
  FILE *f = fopen(path, "wb");
  if (!f) { perror(path); exit(3); }
  /* MH_NODUMP=1: timing runs keep only the header (loop_seconds) */
  int32_t n = getenv("MH_NODUMP") ? 0 : nblk;
  fwrite("MHDUMP1", 1, 8, f);
  fwrite(&real_size, 4, 1, f);
  fwrite(&n, 4, 1, f);
  fwrite(&secs, 8, 1, f);
  for (int i = 0; i < n; i++) {
    fwrite(blk[i].name, 1, 16, f);
    fwrite(&blk[i].nbytes, 8, 1, f);
    fwrite(blk[i].base, 1, (size_t)blk[i].nbytes, f);
  }
  fclose(f);
}

void mhdump_(const char *path, long len) {
  char p[256];
  copy_name(p, sizeof p, path, len);
  dump_to(p, 0.0);
}

static void at_exit_hook(void) { dump_to(exit_path, now() - t_loop0); }

void mhatexit_(const char *path, long len) {
  copy_name(exit_path, sizeof exit_path, path, len);
  t_loop0 = now();
  atexit(at_exit_hook);
}
