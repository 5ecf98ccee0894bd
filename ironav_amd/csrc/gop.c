/* gop.c — picture-type decisions of the encoder driver, on the host: the scene-cut measure and the
 * IDR rule of nervosys/IronAV's `H264Encoder::encode` (crates/ironav-codec/src/h264/enc/mod.rs:700-717)
 * with `LowRes` and `change_threads` (crates/ironav-codec/src/change.rs:9-95).
 *
 * A picture starts an IDR picture when it is the first, when `keyint` pictures have passed since the
 * last IDR picture, when the caller forces one, or at a scene cut: its 8x8 blocks of 4x4-averaged luma
 * are predicted from the previous picture's (each moved up to 4 low-resolution samples) worse than
 * SCENE_CUT (0.8) times their error against their own means. */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/ironav_h264w.h"

#define SCENE_CUT 0.8 /* change.rs:6 */

int iw_lowres(const uint8_t *y, int mb_w, int mb_h, int32_t *out) {
  if (!y || !out || mb_w < 1 || mb_h < 1) return IW_EINVAL;
  /* LowRes::new(mb_w*4, mb_h*4, |x, y| luma[y * pw + x]) on the padded picture (enc/mod.rs:686-692,
     change.rs:18-25): the mean of each 4x4 block, rounded down */
  const int lw = mb_w * 4, lh = mb_h * 4, pw = mb_w * 16;
  for (int by = 0; by < lh; by++)
    for (int bx = 0; bx < lw; bx++) {
      const uint8_t *p = y + (size_t)by * 4 * pw + bx * 4;
      int s = 0;
      for (int k = 0; k < 16; k++) s += p[(k >> 2) * pw + (k & 3)];
      out[by * lw + bx] = s >> 4;
    }
  return IW_OK;
}

typedef struct {
  const int32_t *prev, *cur;
  int lw, lh, row0, row1; /* block rows row0, row0 + 8, ... below row1 (low-resolution samples) */
  int64_t inter, intra;
} ChangeJob;

static void *change_rows(void *arg) {
  ChangeJob *j = (ChangeJob *)arg;
  const int lw = j->lw, lh = j->lh;
  int64_t inter = 0, intra = 0;
  for (int by = j->row0; by < j->row1 && by + 8 <= lh; by += 8)
    for (int bx = 0; bx + 8 <= lw; bx += 8) {
      const int32_t *c = j->cur + (size_t)by * lw + bx;
      int sum = 0;
      for (int k = 0; k < 64; k++) sum += c[(k >> 3) * lw + (k & 7)];
      const int mean = sum / 64;
      for (int k = 0; k < 64; k++) {
        const int d = c[(k >> 3) * lw + (k & 7)] - mean;
        intra += d < 0 ? -d : d;
      }
      int64_t best = INT64_MAX;
      for (int dy = -4; dy <= 4; dy++)
        for (int dx = -4; dx <= 4; dx++) {
          const int ox = bx + dx, oy = by + dy;
          if (ox < 0 || oy < 0 || ox + 8 > lw || oy + 8 > lh) continue;
          const int32_t *p = j->prev + (size_t)oy * lw + ox;
          int64_t sad = 0;
          for (int k = 0; k < 64; k++) {
            const int d = c[(k >> 3) * lw + (k & 7)] - p[(k >> 3) * lw + (k & 7)];
            sad += d < 0 ? -d : d;
          }
          if (sad < best) best = sad;
        }
      inter += best;
    }
  j->inter = inter;
  j->intra = intra;
  return NULL;
}

int iw_change(const int32_t *prev, const int32_t *cur, int lw, int lh, int threads, int64_t *inter, int64_t *intra) {
  if (!prev || !cur || !inter || !intra || lw < 1 || lh < 1) return IW_EINVAL;
  /* block rows 0, 8, ... while a whole 8x8 block fits (change.rs:64); up to `threads` threads take
     consecutive runs of them, and the sums are the same for any count */
  const int nrows = lh >= 8 ? (lh - 8) / 8 + 1 : 0;
  int t = threads < 1 ? 1 : (threads > 64 ? 64 : threads);
  if (t > nrows) t = nrows > 0 ? nrows : 1;
  const int per = nrows > 0 ? (nrows + t - 1) / t : 0;
  ChangeJob jobs[64];
  pthread_t th[64];
  int live[64] = {0};
  for (int k = 0; k < t; k++) {
    const int r0 = k * per, r1 = (k + 1) * per < nrows ? (k + 1) * per : nrows;
    jobs[k] = (ChangeJob){prev, cur, lw, lh, r0 * 8, r1 * 8, 0, 0};
    if (k > 0) live[k] = pthread_create(&th[k], NULL, change_rows, &jobs[k]) == 0;
  }
  change_rows(&jobs[0]);
  int64_t a = jobs[0].inter, b = jobs[0].intra;
  for (int k = 1; k < t; k++) {
    if (live[k]) pthread_join(th[k], NULL);
    else change_rows(&jobs[k]); /* a thread that did not start: its rows on this one */
    a += jobs[k].inter;
    b += jobs[k].intra;
  }
  *inter = a;
  *intra = b;
  return IW_OK;
}

struct iw_gop {
  int mb_w, mb_h, threads;
  uint64_t keyint, count, idr_at;
  int have_last;
  int32_t *last, *cur;
};

int iw_gop_create(iw_gop **out, int width, int height, uint64_t keyint, int threads) {
  if (!out || width < 16 || height < 16 || width > 8192 || height > 8192) return IW_EINVAL;
  iw_gop *g = (iw_gop *)calloc(1, sizeof *g);
  if (!g) return IW_ENOMEM;
  g->mb_w = (width + 15) / 16;
  g->mb_h = (height + 15) / 16;
  g->keyint = keyint ? keyint : 250; /* KEYINT, enc/mod.rs:27; with_keyint clamps to >= 1 (:257-259) */
  g->threads = threads < 1 ? 1 : threads;
  const size_t n = (size_t)g->mb_w * 4 * g->mb_h * 4;
  g->last = (int32_t *)malloc(n * sizeof(int32_t));
  g->cur = (int32_t *)malloc(n * sizeof(int32_t));
  if (!g->last || !g->cur) {
    iw_gop_destroy(g);
    return IW_ENOMEM;
  }
  *out = g;
  return IW_OK;
}

void iw_gop_destroy(iw_gop *g) {
  if (!g) return;
  free(g->last);
  free(g->cur);
  free(g);
}

int iw_gop_next(iw_gop *g, const uint8_t *y, int force_idr, int *kind, int64_t *inter, int64_t *intra) {
  if (!g || !y || !kind) return IW_EINVAL;
  /* enc/mod.rs:706-714 */
  const uint64_t d = g->count++;
  iw_lowres(y, g->mb_w, g->mb_h, g->cur);
  int64_t a = 0, b = 0;
  int cut = 0;
  if (g->have_last) {
    iw_change(g->last, g->cur, g->mb_w * 4, g->mb_h * 4, g->threads, &a, &b);
    cut = (double)a > SCENE_CUT * (double)b;
  }
  int32_t *t = g->last;
  g->last = g->cur;
  g->cur = t;
  g->have_last = 1;
  const int idr = force_idr || d == 0 || d - g->idr_at >= g->keyint || cut;
  if (idr) g->idr_at = d; /* code_picture records the IDR picture's number (enc/mod.rs:543-545) */
  *kind = idr ? 0 : 1;
  if (inter) *inter = a;
  if (intra) *intra = b;
  return IW_OK;
}
