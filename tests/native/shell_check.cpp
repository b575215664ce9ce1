// Host-side check of the sweep planning (test infrastructure): runs the host
// set-up of each mesh against the stand-in device library, replays the
// library's own planning code (fast_boxes, fill_volume_args, build_shell)
// on the captured set-up, and checks every cell of the allocation:
//   - inside a fast box (exact, or cut to whole lines of 16 or 32 elements)
//     every component takes the ordinary update;
//   - every other cell lies in at most one shell piece, whose x runs give
//     the cell's own regions (and UPML slab indices), and a cell in no piece
//     has no component updated;
//   - a piece's rows shorter than half a line are padded to whole lines
//     (lanes per row a multiple of the line's elements, the first cell at
//     its own lane of a line), longer rows are not;
//   - the launch table lists every block of every piece exactly once, each
//     XCD (block b runs on XCD b % 8) taking one contiguous run of blocks of
//     each piece, and the runs of a piece differ in length by at most one.
// Usage: shell_check OUTDIR MESH...   (one line per mesh; exit 1 on the
// first failure)
#include <cstdio>
#include <cstring>
#include "vgpu_run.h"

extern "C" {
// vulture_amd/host/vh_gpu.h (C99 only)
int vh_run_mesh_gpu(const char *mesh, const char *outdir, int prec, int device, int fullPrecision, long steps,
                    unsigned long chunk, double *loop_ms, char *err, int errlen);
extern int vgpu_stub_capture;
extern vgpu_grid_desc vgpu_stub_grid;
extern unsigned vgpu_stub_pmlmask, vgpu_stub_murmask;
extern int vgpu_stub_fplim[6][6][6], vgpu_stub_fmlim[6][6][6];
}

using namespace vgpu;

static int fail(const char *mesh, const char *what, int i, int j, int k, int C)
{
  printf("%s FAIL %s at (%d,%d,%d) component %d\n", mesh, what, i, j, k, C);
  return 1;
}

template <int NB, typename A>
static int check_sweep(const char *mesh, const A &a, const int (*in)[NB][6], const int *fb, const int N[3],
                       long long &npcs, long long &shell, int lw)
{
  std::vector<ShellPiece> pcs;
  std::vector<ShellWork> work;
  std::string err;
  if (build_shell<NB>(a, in, fb, N, pcs, work, shell, err, lw)) {
    printf("%s FAIL build_shell: %s\n", mesh, err.c_str());
    return 1;
  }
  npcs = (long long)pcs.size();
  const long long ncell = (long long)N[0] * N[1] * N[2];
  std::vector<int> owner(ncell, -1);
  for (size_t q = 0; q < pcs.size(); ++q) {
    const ShellPiece &pc = pcs[q];
    if (pc.ni != pc.box[1] - pc.box[0] + 1 || pc.nj != pc.box[3] - pc.box[2] + 1 ||
        (long long)pc.cells != (long long)pc.ni * pc.nj * (pc.box[5] - pc.box[4] + 1))
      return fail(mesh, "piece extents", pc.box[0], pc.box[2], pc.box[4], -1);
    for (int C = 0; C < 3; ++C) {
      if (pc.nrun[C] < 0 || pc.nrun[C] > SHELL_RUNS) return fail(mesh, "run count", pc.box[0], pc.box[2], pc.box[4], C);
      for (int r = 0; r < pc.nrun[C]; ++r) {
        const ShellRun &rn = pc.run[C][r];
        if (rn.xa < pc.box[0] || rn.xb > pc.box[1] || rn.xa > rn.xb || (r > 0 && rn.xa <= pc.run[C][r - 1].xb))
          return fail(mesh, "run outside its piece or out of order", rn.xa, rn.xb, r, C);
      }
    }
    for (int k = pc.box[4]; k <= pc.box[5]; ++k)
      for (int j = pc.box[2]; j <= pc.box[3]; ++j)
        for (int i = pc.box[0]; i <= pc.box[1]; ++i) {
          if (i < 0 || j < 0 || k < 0 || i >= N[0] || j >= N[1] || k >= N[2])
            return fail(mesh, "piece outside the grid", i, j, k, -1);
          int &o = owner[((long long)k * N[1] + j) * N[0] + i];
          if (o >= 0) return fail(mesh, "cell in two pieces", i, j, k, -1);
          o = (int)q;
        }
  }
  const bool hasFast = fb[0] <= fb[1] && fb[2] <= fb[3] && fb[4] <= fb[5];
  for (int k = 0; k < N[2]; ++k)
    for (int j = 0; j < N[1]; ++j)
      for (int i = 0; i < N[0]; ++i) {
        const int o = owner[((long long)k * N[1] + j) * N[0] + i];
        const bool inFast = hasFast && in_box(fb, i, j, k);
        if (inFast && o >= 0) return fail(mesh, "fast cell in a piece", i, j, k, -1);
        for (int C = 0; C < 3; ++C) {
          long long p = -1;
          const int reg = classify<NB>(a, in[C], C, i, j, k, p);
          if (inFast) {
            if (reg != 6) return fail(mesh, "fast cell not ordinary", i, j, k, C);
          } else if (o < 0) {
            if (reg != -1) return fail(mesh, "updated cell in no piece", i, j, k, C);
          } else {
            const ShellPiece &pc = pcs[o];
            const int dj = j - pc.box[2], dk = k - pc.box[4];
            // the kernels take the region at the first plane of their z chunk
            // and step the slab index by psz per plane
            const int dk0 = dk / pc.kc * pc.kc;
            const ColRegion rg = C == 0 ? shell_region<0>(&pc, i, dj, dk0)
                               : C == 1 ? shell_region<1>(&pc, i, dj, dk0)
                                        : shell_region<2>(&pc, i, dj, dk0);
            if (rg.reg != reg) return fail(mesh, "piece region differs from the cell's", i, j, k, C);
            if (reg >= 0 && reg < 6 && rg.p + (dk - dk0) * rg.psz != p)
              return fail(mesh, "UPML slab index", i, j, k, C);
          }
        }
      }
  // launch table
  if (work.size() % 8) return fail(mesh, "launch table not a multiple of 8", 0, 0, 0, -1);
  std::vector<std::vector<int>> seen(pcs.size());
  for (size_t q = 0; q < pcs.size(); ++q) {
    const ShellPiece &pc = pcs[q];
    // rows shorter than half a line: nip lanes, a whole number of lw-element
    // lines, the first cell at lane xo = box[0] % lw; other rows unpadded
    const bool rows = lw && 2 * pc.ni < lw
                          ? pc.xo == pc.box[0] % lw && pc.nip % lw == 0 && pc.nip >= pc.xo + pc.ni &&
                                pc.nip < pc.xo + pc.ni + lw
                          : pc.xo == 0 && pc.nip == pc.ni;
    if (!rows || pc.nk != pc.box[5] - pc.box[4] + 1 || pc.nxy != pc.nip * pc.nj || pc.nxyb != (pc.nxy + 255) / 256 ||
        pc.kc < 1 || pc.kc > pc.nk || pc.nblk != pc.nxyb * ((pc.nk + pc.kc - 1) / pc.kc))
      return fail(mesh, "piece block layout", pc.box[0], pc.box[2], pc.box[4], -1);
    seen[q].assign(pc.nblk, -1);
  }
  for (size_t b = 0; b < work.size(); ++b) {
    const ShellWork w = work[b];
    if (w.piece < 0) continue;
    if (w.piece >= (int)pcs.size() || w.block < 0 || w.block >= (int)seen[w.piece].size())
      return fail(mesh, "launch entry out of range", (int)b, w.piece, w.block, -1);
    if (seen[w.piece][w.block] >= 0) return fail(mesh, "block launched twice", (int)b, w.piece, w.block, -1);
    seen[w.piece][w.block] = (int)(b % 8);
  }
  for (size_t q = 0; q < pcs.size(); ++q) {
    const int nb = (int)seen[q].size();
    int runs = 0, lo = nb, hi = 0;
    for (int b = 0; b < nb; ++b) {
      if (seen[q][b] < 0) return fail(mesh, "block never launched", (int)q, b, 0, -1);
      if (b > 0 && seen[q][b] < seen[q][b - 1]) return fail(mesh, "XCD runs out of order", (int)q, b, 0, -1);
    }
    for (int x = 0; x < 8; ++x) {
      int n = 0;
      for (int b = 0; b < nb; ++b) n += seen[q][b] == x;
      lo = std::min(lo, n);
      hi = std::max(hi, n);
      runs += n > 0;
    }
    if (hi - lo > 1) return fail(mesh, "unbalanced XCD runs", (int)q, lo, hi, -1);
    (void)runs;
  }
  return 0;
}

// The fused sweep's tiling (k_leap): tile tx owns x-cells box[0] + 63 tx +
// lane for lanes 0..62, wavefront w of y-block ty owns rows box[2] +
// (4 ty + w) R + r for r < R, z chunk tz owns planes box[4] + kz tz + p for
// p < kz, each clipped to the box.  Every cell of the box must be owned
// exactly once, for each row count R and chunk depth kz the library uses.
static int check_leap_tiles(const char *mesh, const int *fb, const int N[3])
{
  if (fb[0] > fb[1] || fb[2] > fb[3] || fb[4] > fb[5]) return 0;
  if (fb[0] < 1 || fb[2] < 1 || fb[4] < 1 || fb[1] > N[0] - 2 || fb[3] > N[1] - 2 || fb[5] > N[2] - 2)
    return fail(mesh, "fused box not inside the grid", fb[0], fb[2], fb[4], -1);
  const int nx = fb[1] - fb[0] + 1, ny = fb[3] - fb[2] + 1, nz = fb[5] - fb[4] + 1;
  for (int R : {2, 4})
    for (int kz0 : {1, 8, 16}) {
      const int kz = std::min(kz0, nz);
      const int ntx = (nx + 62) / 63, ngy = (ny + 4 * R - 1) / (4 * R), ntz = (nz + kz - 1) / kz;
      std::vector<unsigned char> seen((size_t)nx * ny * nz, 0);
      for (int tz = 0; tz < ntz; ++tz)
        for (int ty = 0; ty < ngy; ++ty)
          for (int w = 0; w < 4; ++w)
            for (int tx = 0; tx < ntx; ++tx) {
              const int jb = fb[2] + (ty * 4 + w) * R, kb = fb[4] + tz * kz;
              if (jb > fb[3] || kb > fb[5]) continue;
              const int ke = std::min(kb + kz, fb[5] + 1);
              for (int k = kb; k < ke; ++k)
                for (int r = 0; r < R && jb + r <= fb[3]; ++r)
                  for (int lane = 0; lane < 63; ++lane) {
                    const int i = fb[0] + 63 * tx + lane;
                    if (i > fb[1]) break;
                    unsigned char &s = seen[((size_t)(k - fb[4]) * ny + (jb + r - fb[2])) * nx + (i - fb[0])];
                    if (s) return fail(mesh, "fused tile cell owned twice", i, jb + r, k, R);
                    s = 1;
                  }
            }
      for (size_t q = 0; q < seen.size(); ++q)
        if (!seen[q]) return fail(mesh, "fused tile cell not owned", (int)(q % nx), (int)(q / nx % ny), (int)(q / nx / ny), R);
    }
  return 0;
}

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: shell_check OUTDIR MESH...\n");
    return 2;
  }
  vgpu_stub_capture = 1;
  for (int m = 2; m < argc; ++m) {
    char err[512] = "";
    double ms = 0;
    if (vh_run_mesh_gpu(argv[m], argv[1], 0, 0, 1, 1, 0, &ms, err, sizeof err) != 0) {
      printf("%s FAIL set-up: %s\n", argv[m], err);
      return 1;
    }
    Ctx *c = new Ctx();
    c->grid = vgpu_stub_grid;
    c->NX = c->grid.nx;
    c->NY = c->grid.ny;
    c->NZ = c->grid.nz;
    c->pmlmask = vgpu_stub_pmlmask;
    c->murmask = vgpu_stub_murmask;
    memcpy(c->fplim, vgpu_stub_fplim, sizeof(c->fplim));
    memcpy(c->fmlim, vgpu_stub_fmlim, sizeof(c->fmlim));
    const int N[3] = {c->NX, c->NY, c->NZ};
    EArgs<double> ea{};
    HArgs<double> ha{};
    fill_volume_args<double>(c, ea, ha, 0);
    int ein[3][1][6];
    for (int C = 0; C < 3; ++C)
      for (int q = 0; q < 6; ++q) ein[C][0][q] = ea.in[C][q];
    long long pe = 0, ph = 0, se = 0, sh = 0;
    // exact fast boxes, boxes cut to whole lines of FP64 and FP32 elements,
    // and the fused box of the leapfrog sweep (the same box for E and H)
    for (int lw : {0, 16, 32, -1}) {
      if (lw >= 0) {
        fast_boxes(c->grid.gfilim, N, c->fastE, c->fastH, lw);
      } else {
        leap_box(c->grid.gfilim, N, c->fastE);
        memcpy(c->fastH, c->fastE, sizeof(c->fastH));
        if (check_leap_tiles(argv[m], c->fastE, N)) return 1;
      }
      const int rw = lw < 0 ? 32 : lw;
      if (check_sweep<1>(argv[m], ea, ein, c->fastE, N, pe, se, rw) ||
          check_sweep<3>(argv[m], ha, ha.in, c->fastH, N, ph, sh, rw))
        return 1;
    }
    printf("%s OK cells %lld E pieces %lld shell %lld H pieces %lld shell %lld\n", argv[m],
           (long long)N[0] * N[1] * N[2], pe, se, ph, sh);
    delete c;
  }
  return 0;
}
