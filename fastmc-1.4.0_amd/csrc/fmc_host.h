// Host-side set-up of the trial-move path: cell algebra, tables, k-vector
// list, tail corrections.  Pure C++ (no HIP types); used by the single-chain
// context and by the batched chains.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fmc {

constexpr double kR4pie0 = 138935.4835;   // utility_pack.f:95
constexpr double kPi = 3.141592653589793;  // utility_pack.f:74
constexpr double kSqrpi = 1.7724538509055159;
constexpr double kKboltz = 1.9872041e-3;   // kcal/mol/K, utility_pack.f:103
constexpr double kBoltz = 1.3806503e-23;   // J/K, utility_pack.f:107
constexpr int kMaxSites = 8;               // sites per rigid guest on device
constexpr int kMaxGuestTypes = 4;

void dcell(const double a[9], double b[10]);
void invert(const double a[9], double b[9], double* det);
double volume(int imcon, const double cell[9]);
void ewald_precision(double eps, double rcut, int imcon, const double cell[9],
                     double* alpha, int kmax[3]);
int mxgrid_for(double rvdw);

// One k-vector of the half-space enumeration (ewald.f:141-213): integer
// indices and the Gaussian weight a_k = exp(-k^2/4 alpha^2)/k^2 (x4 for
// imcon 4,5,7).
struct KVec {
  int l, m, n;
  double ak;
};
// k-vectors in the reference's enumeration order
std::vector<KVec> enumerate_kvecs(int imcon, const double cell[9],
                                  double alpha, const int kmax[3]);

struct GuestTemplate {
  int ns = 0;
  double x[kMaxSites], y[kMaxSites], z[kMaxSites];
  double q[kMaxSites], mass[kMaxSites];
  int type[kMaxSites];
  double e3 = 0.0;     // ewald3 exclusion constant (gcmc.f:529-552)
  double engsic = 0.0; // -r4pie0/epsq * alpha * sum q^2 / sqrt(pi)
};

// Everything static about one chain's system.
struct SystemDesc {
  int imcon = 0;
  double cell[9] = {0};
  double rcell[9] = {0};
  double volm = 0;
  double alpha = 0, epsq = 1.0;
  int kmax[3] = {0, 0, 0};
  bool have_ewald = false;
  double rcut = 0, delr = 1.0, rvdw = 0;
  int ntpatm = 0, ntpvdw = 0;
  std::vector<int> vdw_a, vdw_b, vdw_key;
  std::vector<double> vdw_prm;  // 5 per potential, internal units
  // framework
  std::vector<double> fx, fy, fz, fq;
  std::vector<int> ftype, ffrz;
  std::vector<GuestTemplate> guests;
  // initial guest molecules per type: coordinates [mol*ns + s]
  std::vector<std::vector<double>> gx0, gy0, gz0;
  int max_guest_atoms = 3000;
  // derived at finalize
  int mxgrid = 0;
  double drewd = 0, dlrpot = 0;
  std::vector<double> erc;        // [0..mxgrid], [0] duplicates [1]
  std::vector<double> vvv;        // [t*(mxgrid+1) + i], t = 0..ntable-1
  int ntable = 0;                 // tables incl. the zero table
  std::vector<int> lstvdw;        // loc8-1 -> 1-based table index
  std::vector<int> ltpvdw;        // per table (0-based): key, -1 none
  std::vector<int> pairtab;       // [ti*ntpatm+tj] (0-based types) -> table or -1
  std::vector<double> steadd;     // per loc8 pair
  std::vector<KVec> kvecs;
  double fw_engsic = 0;
  bool finalized = false;

  int finalize(std::string* err);
  // gstlrcorrect-style tail for given per-type counts (vdw_terms.f:684-745)
  double tail(const std::vector<int>& numtyp,
              const std::vector<int>& numfrz) const;
  void counts(const std::vector<int>& nmol, std::vector<int>* numtyp,
              std::vector<int>* numfrz) const;
};

// ewald3 over one molecule's actual coordinates (exclusion correction of
// ewald.f:697-758 for every intramolecular pair)
double ewald3_molecule(const SystemDesc& s, const GuestTemplate& g,
                       const double* x, const double* y, const double* z);

}  // namespace fmc
