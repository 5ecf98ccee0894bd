// analyse_host.cpp — the analysis kernel (ironav_amd/csrc/analyse.hip) built for the host with
// hostwave.h's lane emulation, driven like ia_analyse_picture.
//
// TEST INFRASTRUCTURE. Usage: analyse_host IN OUT. IN holds int32 fields (width, height, cabac,
// transform8, sub8x8, fast, slices, search_mode, qp, kind, nref, has_mbs0, has_fs), then the padded
// planes y, cb, cr (u8), per reference its four half-sample planes and its cb, cr planes, then
// reference 0's ia_mbinfo records if has_mbs0, then ia_fsresult[nref][mb][IA_NPART] if has_fs.
// OUT receives ia_mbcode[mb], ia_mbstats[mb], the reconstruction y, cb, cr and ia_mbinfo[mb].
// The parameters are derived as the library's analysis_params derives them (ironav_amd/csrc/api.hip).
#include "../../ironav_amd/csrc/analyse.hip"

static void *take(FILE *f, size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p || fread(p, 1, n, f) != n) {
    fprintf(stderr, "analyse_host: short input\n");
    exit(2);
  }
  return p;
}

static int chroma_qp(int qp) {
  static const uint8_t T[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
  return qp < 30 ? qp : T[qp - 30];
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: analyse_host IN OUT\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t h[15];
  if (fread(h, sizeof h, 1, f) != 1) return 2;
  const int width = h[0], height = h[1], cabac = h[2], t8 = h[3], sub8 = h[4], fast = h[5], slices = h[6];
  const int qp = h[8], kind = h[9], nref = h[10], has_mbs0 = h[11], has_fs = h[12], tb = h[13], td = h[14];
  const ia::Geom g = ia::make_geom(width, height);
  const size_t ylen = (size_t)g.pw * g.ph, clen = (size_t)g.cw * g.ch, hplen = 4 * (size_t)g.sw * g.sh;
  const size_t nmb = (size_t)g.mb_w * g.mb_h;
  ia::AnParams p;
  memset(&p, 0, sizeof p);
  for (int k = 0; k < 3; k++) p.src[k] = (const uint8_t *)take(f, k ? clen : ylen);
  for (int r = 0; r < nref; r++) {
    p.hp[r] = (const uint8_t *)take(f, hplen);
    p.rcb[r] = (const uint8_t *)take(f, clen);
    p.rcr[r] = (const uint8_t *)take(f, clen);
  }
  if (has_mbs0) p.rmbs0 = (const ia_mbinfo *)take(f, nmb * sizeof(ia_mbinfo));
  if (has_fs) p.fs = (const ia_fsresult *)take(f, (size_t)nref * nmb * IA_NPART * sizeof(ia_fsresult));
  fclose(f);
  for (int k = 0; k < 3; k++) p.rec[k] = (uint8_t *)calloc(k ? clen : ylen, 1);
  p.codes = (ia_mbcode *)calloc(nmb, sizeof(ia_mbcode));
  p.stats = (ia_mbstats *)calloc(nmb, sizeof(ia_mbstats));
  p.mbs = (ia_mbinfo *)calloc(nmb, sizeof(ia_mbinfo));
  p.state = (ia::MbState *)calloc(nmb, sizeof(ia::MbState));
  uint32_t *ent = (uint32_t *)malloc(128 * sizeof(uint32_t));
  for (int s = 0; s < 64; s++) {
    const double lps = 0.5 * pow(0.01875 / 0.5, (double)s / 63.0);
    ent[2 * s] = (uint32_t)round(-log2(1.0 - lps) * 256.0);
    ent[2 * s + 1] = (uint32_t)round(-log2(lps) * 256.0);
  }
  p.entropy = ent;
  p.mb_w = g.mb_w;
  p.mb_h = g.mb_h;
  p.nslices = slices < g.mb_h ? slices : g.mb_h;
  // a B picture (kind 2): the list 0 and list 1 pictures, the list 1 picture's records
  p.nref = kind == IA_PIC_P ? nref : kind == IA_PIC_B ? 1 : 0;
  p.b_slice = kind == IA_PIC_B;
  p.dist_tb = tb;
  p.dist_td = td;
  p.qp_slice = qp;
  p.qp = qp;
  p.qpc = chroma_qp(qp);
  p.lambda = (int)round(pow(2.0, ((double)qp - 12.0) / 6.0));
  if (p.lambda < 1) p.lambda = 1;
  p.p_slice = kind == IA_PIC_P || kind == IA_PIC_B;
  p.sub8 = sub8 && qp <= 20;
  p.use_cabac = cabac != 0;
  p.transform8 = cabac && t8;
  p.lam2 = 1.6 * pow(2.0, ((double)qp - 12.0) / 3.0);
  p.skip_thresh = 256.0 * 0.625 * pow(2.0, (double)qp / 6.0);
  const double step = pow(2.0, (double)qp / 6.0);
  p.fast = fast != 0;
  p.fast_skip_y = 256.0 * 0.05 * step;
  p.fast_skip_c = 128.0 * 0.05 * step;
  p.fast_flat = 256.0 * 0.8 * step;
  if (ia::launch_analyse(p, nullptr) != hipSuccess) return 3;
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(p.codes, sizeof(ia_mbcode), nmb, o);
  fwrite(p.stats, sizeof(ia_mbstats), nmb, o);
  fwrite(p.rec[0], 1, ylen, o);
  fwrite(p.rec[1], 1, clen, o);
  fwrite(p.rec[2], 1, clen, o);
  fwrite(p.mbs, sizeof(ia_mbinfo), nmb, o);
  fclose(o);
  return 0;
}
