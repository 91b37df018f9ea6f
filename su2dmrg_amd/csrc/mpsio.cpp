// Checkpoint files of the DMRG driver in the reference's format.
//
// uu-{ns}-{D}-{2S}-{ky}-{exci}-{i}.dat / .bin (dmrg_su2::save_mps2 and
// read_mps, dmrg_su2_omp.cpp:864-1068): the text header lists nbond and
// locspin, then per leg its direction, sector count, 2j labels and multiplet
// counts (tab-terminated lines); the binary body holds one int32 flag per
// block index (1: present, tensor_su2.cpp:566-583), the reduced doubles of
// the present blocks, then their Clebsch-Gordan doubles, both in block-index
// order (leg 0 fastest) and column-major inside a block.
//
// Site i < ns/2 is the three-leg uu[i] = [lb[i-1], site, lb[i]] with CG
// make_cgc(j_a, 1, j_b); site i >= ns/2 is [site, rb[i+1], rb[i]] with
// make_cgc(1, j_r', j_r) -- the tensors tensor_su2::svd leaves
// (tensor_su2.cpp:1733-1743).  Block (j_a, 1, j_b) holds the projector rows of
// the part j_a inside the fused sector j_b.  ww-*.bin (save_ww, 1154-1186)
// holds max_dcut * phdim doubles of the last solve's singular values, zero
// beyond them; enr-*.dat (save_enr, 1070-1098) gets one line, the Ritz value
// minus 1000 at 16 significant digits.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

#include "cg.hpp"
#include "common.hpp"
#include "context.hpp"
#include "dmrg.hpp"
#include "mpsio.hpp"
#include "ops.hpp"

namespace su2 {

namespace mpsio {


std::string uu_base(const std::string& dir, int ns, int d, int s, int ky, int exci) {
  char b[64];
  std::snprintf(b, sizeof b, "uu-%d-%d-%d-%d-%d-", ns, d, s, ky, exci);
  return dir + "/" + b;
}

std::string tag(int ns, int d, int s, int ky, int exci) {
  char b[64];
  std::snprintf(b, sizeof b, "%d-%d-%d-%d-%d", ns, d, s, ky, exci);
  return b;
}

long long block_count(const Leg* l) { return (long long)l[0].nsec() * l[1].nsec() * l[2].nsec(); }

void labels_of(const Leg* l, long long idx, int* s, int* tj) {
  s[0] = (int)(idx % l[0].nsec());
  s[1] = (int)((idx / l[0].nsec()) % l[1].nsec());
  s[2] = (int)(idx / ((long long)l[0].nsec() * l[1].nsec()));
  for (int k = 0; k < 3; ++k) tj[k] = l[k].twoj[s[k]];
}

void write_file(const std::string& path, const File3& f) {
  std::ofstream dat(path + ".dat");
  SU2_CHECK(dat.is_open(), kErrArgument, "save: cannot write %s.dat", path.c_str());
  dat << 3 << "\t" << 0 << "\n";
  for (int k = 0; k < 3; ++k) {
    dat << f.leg[k].dir << "\t" << f.leg[k].nsec() << "\n";
    for (int j : f.leg[k].twoj) dat << j << "\t";
    dat << "\n";
    for (int d : f.leg[k].dim) dat << d << "\t";
    dat << "\n";
  }
  std::ofstream bin(path + ".bin", std::ios::binary);
  SU2_CHECK(bin.is_open(), kErrArgument, "save: cannot write %s.bin", path.c_str());
  bin.write((const char*)f.flag.data(), f.flag.size() * sizeof(int));
  bin.write((const char*)f.red.data(), f.red.size() * sizeof(double));
  bin.write((const char*)f.cg.data(), f.cg.size() * sizeof(double));
}

File3 read_file(const std::string& path) {
  File3 f;
  std::ifstream dat(path + ".dat");
  SU2_CHECK(dat.is_open(), kErrArgument, "load: cannot read %s.dat", path.c_str());
  int nb = 0, loc = 0;
  dat >> nb >> loc;
  SU2_CHECK(nb == 3 && loc == 0, kErrStructure, "load: %s.dat is not a three-leg MPS tensor",
            path.c_str());
  for (int k = 0; k < 3; ++k) {
    int dir = 0, n = 0;
    dat >> dir >> n;
    SU2_CHECK(dat.good() && n >= 1 && n <= 64, kErrStructure, "load: bad header in %s.dat",
              path.c_str());
    std::vector<int> tj(n), dm(n);
    for (int& x : tj) dat >> x;
    for (int& x : dm) dat >> x;
    SU2_CHECK(!dat.fail(), kErrStructure, "load: truncated header in %s.dat", path.c_str());
    f.leg[k] = make_leg(tj, dm, dir);
  }
  std::ifstream bin(path + ".bin", std::ios::binary);
  SU2_CHECK(bin.is_open(), kErrArgument, "load: cannot read %s.bin", path.c_str());
  const long long nten = block_count(f.leg);
  f.flag.resize(nten);
  bin.read((char*)f.flag.data(), nten * sizeof(int));
  long long nred = 0, ncg = 0;
  int s[3], tj[3];
  for (long long i = 0; i < nten; ++i) {
    if (!f.flag[i]) continue;
    labels_of(f.leg, i, s, tj);
    SU2_CHECK(labels_allowed(tj, 3, 0), kErrStructure, "load: %s.bin flags a forbidden block",
              path.c_str());
    nred += (long long)f.leg[0].dim[s[0]] * f.leg[1].dim[s[1]] * f.leg[2].dim[s[2]];
    ncg += (long long)(tj[0] + 1) * (tj[1] + 1) * (tj[2] + 1);
  }
  f.red.resize(nred);
  f.cg.resize(ncg);
  bin.read((char*)f.red.data(), nred * sizeof(double));
  bin.read((char*)f.cg.data(), ncg * sizeof(double));
  SU2_CHECK(!bin.fail(), kErrStructure, "load: %s.bin is shorter than its header says",
            path.c_str());
  return f;
}

// r with stored = r * ref (the ratio test of tensor::is_proportional_to,
// tensor.cpp:1231-1269, at the largest entry of ref)
double cg_ratio(const double* stored, const std::vector<double>& ref, const std::string& what) {
  size_t pos = 0;
  for (size_t i = 1; i < ref.size(); ++i)
    if (std::fabs(ref[i]) > std::fabs(ref[pos])) pos = i;
  const double r = stored[pos] / ref[pos];
  for (size_t i = 0; i < ref.size(); ++i)
    SU2_CHECK(std::fabs(stored[i] - r * ref[i]) <= 1e-10 * (1 + std::fabs(r)), kErrStructure,
              "load: %s carries a CG that is not a multiple of make_cgc", what.c_str());
  return r;
}

double read_energy(const std::string& dir, const std::string& t) {
  std::ifstream f(dir + "/enr-" + t + ".dat");
  SU2_CHECK(f.is_open(), kErrArgument, "load: no enr-%s.dat in %s", t.c_str(), dir.c_str());
  double v = 0, energy = 0;
  bool any = false;
  while (f >> v) {
    energy = v;
    any = true;
  }
  SU2_CHECK(any, kErrStructure, "load: empty enr-%s.dat in %s", t.c_str(), dir.c_str());
  return energy + 1000.0;
}

}  // namespace mpsio

namespace {

using namespace mpsio;
constexpr int kPhysSpin = 1;

std::vector<double> download(Context& c, const Tensor& t) {
  std::vector<double> h(t.size());
  if (t.size())
    SU2_HIP(hipMemcpyAsync(h.data(), t.data(), t.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c.stream()));
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return h;
}

// three-leg file of projector P (fused leg F.fused -> bond); `block_leg` is
// the leg of the enlarged block (0: left site, 1: right site)
File3 to_file(Context& c, const Tensor& P, const Fused& F, const Leg& bond, bool left) {
  File3 f;
  const Leg site = make_leg({kPhysSpin}, {1}, 1);
  Leg blk = F.block, b = bond;
  blk.dir = 1;
  b.dir = -1;
  if (left) {
    f.leg[0] = blk;
    f.leg[1] = site;
  } else {
    f.leg[0] = site;
    f.leg[1] = blk;
  }
  f.leg[2] = b;
  const int bl = left ? 0 : 1;
  const std::vector<double> h = download(c, P);
  const long long nten = block_count(f.leg);
  f.flag.assign(nten, 0);
  int s[3], tj[3];
  for (long long i = 0; i < nten; ++i) {
    labels_of(f.leg, i, s, tj);
    if (!labels_allowed(tj, 3, 0)) continue;
    const int pb = P.find2(tj[2], tj[2]);
    auto it = F.off.find({tj[bl], tj[2]});
    if (pb < 0 || it == F.off.end()) continue;
    f.flag[i] = 1;
    const Block& B = P.blocks()[pb];
    const int rows = f.leg[bl].dim[s[bl]], cols = B.dim[1];
    for (int col = 0; col < cols; ++col)
      for (int r = 0; r < rows; ++r)
        f.red.push_back(h[B.offset + it->second + r + (long long)col * B.dim[0]]);
    const std::vector<double> g = cg_tensor(tj[0], tj[1], tj[2]);
    f.cg.insert(f.cg.end(), g.begin(), g.end());
  }
  return f;
}

// projector [F.fused, bond] from a three-leg file (block leg bl)
std::unique_ptr<Tensor> from_file(Context& c, const File3& f, const Fused& F, int bl,
                                  const std::string& what) {
  Leg bond = f.leg[2];
  bond.dir = -1;
  Leg fl = F.fused;
  fl.dir = 1;
  std::unique_ptr<Tensor> P(new Tensor(&c, {fl, bond}, 0, nullptr));
  std::vector<double> h(P->size(), 0.0);
  const long long nten = block_count(f.leg);
  long long o = 0, oc = 0;
  int s[3], tj[3];
  for (long long i = 0; i < nten; ++i) {
    if (!f.flag[i]) continue;
    labels_of(f.leg, i, s, tj);
    const int rows = f.leg[bl].dim[s[bl]], cols = f.leg[2].dim[s[2]];
    const std::vector<double> g = cg_tensor(tj[0], tj[1], tj[2]);
    const double r = cg_ratio(f.cg.data() + oc, g, what);
    const int pb = P->find2(tj[2], tj[2]);
    auto it = F.off.find({tj[bl], tj[2]});
    SU2_CHECK(pb >= 0 && it != F.off.end(), kErrStructure,
              "load: %s block (%d,%d,%d) does not fit the chain", what.c_str(), tj[0], tj[1],
              tj[2]);
    const Block& B = P->blocks()[pb];
    for (int col = 0; col < cols; ++col)
      for (int rr = 0; rr < rows; ++rr)
        h[B.offset + it->second + rr + (long long)col * B.dim[0]] =
            r * f.red[o + rr + (long long)col * rows];
    o += (long long)rows * cols;
    oc += (long long)g.size();
  }
  if (!h.empty())
    SU2_HIP(hipMemcpyAsync(P->data(), h.data(), h.size() * sizeof(double),
                           hipMemcpyHostToDevice, c.stream()));
  SU2_HIP(hipStreamSynchronize(c.stream()));
  return P;
}

bool same_sectors(const Leg& a, const Leg& b) { return a.twoj == b.twoj && a.dim == b.dim; }

}  // namespace

void Dmrg::save(const std::string& dir, int ky, int exci) const {
  Context& c = c_;
  c.make_current();
  const int ns = lat_.ns;
  SU2_CHECK(!energies_.empty() && U_.count(ns / 2 - 1), kErrArgument,
            "save: run the warm-up (and sweeps) first");
  const std::string base = uu_base(dir, ns, D_, S_, ky, exci);
  for (int i = 0; i < ns; ++i) {
    const bool left = i < ns / 2;
    if (left) {
      SU2_CHECK(U_.count(i) && lb_.count(i), kErrStructure, "save: no projector at site %d", i);
      write_file(base + std::to_string(i),
                 to_file(c, *U_.at(i), make_fused(lb_.at(i - 1), true), lb_.at(i), true));
    } else {
      SU2_CHECK(V_.count(i) && rb_.count(i), kErrStructure, "save: no projector at site %d", i);
      write_file(base + std::to_string(i),
                 to_file(c, *V_.at(i), make_fused(rb_.at(i + 1), false), rb_.at(i), false));
    }
  }
  std::vector<double> ww((size_t)D_ * (kPhysSpin + 1), 0.0);
  for (size_t k = 0; k < wout_.size() && k < ww.size(); ++k) ww[k] = wout_[k];
  {
    std::ofstream f(dir + "/ww-" + tag(ns, D_, S_, ky, exci) + "-" + std::to_string(ns / 2) +
                        ".bin",
                    std::ios::binary);
    SU2_CHECK(f.is_open(), kErrArgument, "save: cannot write the ww file in %s", dir.c_str());
    f.write((const char*)ww.data(), ww.size() * sizeof(double));
  }
  {
    std::ofstream f(dir + "/enr-" + tag(ns, D_, S_, ky, exci) + ".dat", std::ios::app);
    SU2_CHECK(f.is_open(), kErrArgument, "save: cannot write the enr file in %s", dir.c_str());
    char line[64];
    std::snprintf(line, sizeof line, "%.16g\n", energies_.back() - 1000.0);
    f << line;
  }
}

double Dmrg::load(const std::string& dir, int file_dcut, int ky, int exci) {
  Context& c = c_;
  c.make_current();
  const int ns = lat_.ns;
  const int fd = file_dcut > 0 ? file_dcut : D_;
  const std::string base = uu_base(dir, ns, fd, S_, ky, exci);
  // read everything first: a missing or malformed file leaves the run intact
  std::vector<File3> files(ns);
  for (int i = 0; i < ns; ++i) files[i] = read_file(base + std::to_string(i));
  const double energy = read_energy(dir, tag(ns, fd, S_, ky, exci));
  // decode and check every projector against the chain before touching the run
  std::map<int, Leg> lb, rb;
  lb[-1] = make_leg({0}, {1}, 1);
  rb[ns] = make_leg({S_}, {1}, 1);
  std::vector<std::unique_ptr<Tensor>> P(ns);
  for (int i = 0; i < ns / 2; ++i) {
    const File3& f = files[i];
    const std::string what = "site " + std::to_string(i);
    SU2_CHECK(f.leg[1].nsec() == 1 && f.leg[1].twoj[0] == kPhysSpin, kErrStructure,
              "load: %s is not a left-site tensor", what.c_str());
    SU2_CHECK(same_sectors(f.leg[0], lb.at(i - 1)), kErrStructure,
              "load: %s does not continue the left bond", what.c_str());
    P[i] = from_file(c, f, make_fused(lb.at(i - 1), true), 0, what);
    lb[i] = f.leg[2];
    lb[i].dir = -1;
  }
  for (int i = ns - 1; i >= ns / 2; --i) {
    const File3& f = files[i];
    const std::string what = "site " + std::to_string(i);
    SU2_CHECK(f.leg[0].nsec() == 1 && f.leg[0].twoj[0] == kPhysSpin, kErrStructure,
              "load: %s is not a right-site tensor", what.c_str());
    SU2_CHECK(same_sectors(f.leg[1], rb.at(i + 1)), kErrStructure,
              "load: %s does not continue the right bond", what.c_str());
    P[i] = from_file(c, f, make_fused(rb.at(i + 1), false), 1, what);
    rb[i] = f.leg[2];
    rb[i].dir = -1;
  }
  SU2_CHECK(same_sectors(lb.at(ns / 2 - 1), rb.at(ns / 2)), kErrStructure,
            "load: the two halves meet on different centre bonds");
  // prepare_sweep: environments from both edges
  L_.clear();
  R_.clear();
  U_.clear();
  V_.clear();
  lb_ = lb;
  rb_ = rb;
  L_[-1].reset(new Env());
  L_[-1]->bond = lb_[-1];
  R_[ns].reset(new Env());
  R_[ns]->bond = rb_[ns];
  for (int i = 0; i < ns / 2; ++i) {
    if (i < ns / 2 - 1) L_[i] = project(*P[i], *enlarged_left(i, fused_left(i)));
    U_[i] = std::move(P[i]);
  }
  for (int i = ns - 1; i >= ns / 2; --i) {
    R_[i] = project(*P[i], *enlarged_right(i, fused_right(i)));
    V_[i] = std::move(P[i]);
  }
  energies_.clear();
  history_.clear();
  wout_.clear();
  for (auto& st : lower_) {
    st->OL.erase(st->OL.upper_bound(-1), st->OL.end());
    st->OR.erase(st->OR.begin(), st->OR.lower_bound(ns));
  }
  rebuild_overlaps();
  return energy;
}

}  // namespace su2
