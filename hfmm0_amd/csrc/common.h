// Internal declarations shared by the host-side C++ and the HIP kernels of libhfmm.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace hfmm {

constexpr int kMaxLevels = 24;     // deepest level supported (Morton codes use 3 bits per level)
constexpr int kMaxP = 40;          // largest truncation degree supported

// Error carried across the C-ABI as a status code plus a thread-local message.
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
void set_last_error(const std::string& msg);   // message returned by hfmm_last_error()

// ---------------------------------------------------------------------------------
// Octree (canonical order: level ascending, Morton code ascending)
// ---------------------------------------------------------------------------------
struct Tree {
    int depth = 0;                       // finest level
    double pmin[3] = {0, 0, 0};
    double extent = 1.0;                 // side of the root cube
    int64_t n_points = 0;
    int maxleaf = 0;
    std::vector<int> level;              // per box
    std::vector<uint64_t> code;          // Morton code at the box's own level
    std::vector<int> count;              // points in the box's subtree
    std::vector<int> pstart;             // first sorted point of the subtree
    std::vector<int> parent;             // -1 at level 1
    std::vector<uint8_t> is_leaf;
    std::vector<int> level_start;        // size depth+2; boxes of level l: [level_start[l], level_start[l+1])
    std::vector<int64_t> perm;           // sorted index -> original index
    int n_boxes() const { return (int)level.size(); }
    int find(int lvl, uint64_t c) const; // canonical index or -1
};

Tree build_tree(const double* xyz, int64_t n, int n_max, int max_depth /* <=0: auto */);

struct Lists {
    // all as (target, source) box pairs sorted by (target, source)
    std::vector<int> v_tgt, v_src;
    std::vector<int> u_tgt, u_src;       // adaptive U list without self pairs
    std::vector<int> w_tgt, w_src;
    std::vector<int> x_tgt, x_src;
};
Lists build_lists(const Tree& t);
// The same tree and lists built on the current GPU (identical arrays).
void build_tree_device(const double* xyz, int64_t n, int n_max, int max_depth, Tree& t, Lists& ls);

void morton_grid(uint64_t code, int& x, int& y, int& z);
uint64_t morton3(uint64_t x, uint64_t y, uint64_t z);

// ---------------------------------------------------------------------------------
// Operator tables (host, FP64)
// ---------------------------------------------------------------------------------
inline int nm_index(int n, int m) { return n * n + n + m; }
inline int ncoef(int p) { return (p + 1) * (p + 1); }
// offset of degree n's (2n+1)^2 block in a packed rotation table
inline int rot_offset(int n) { return n * (2 * n - 1) * (2 * n + 1) / 3; }
inline int rot_size(int p) { return rot_offset(p + 1); }
// packed coaxial operator: for m = 0..min(pt,ps), a (pt+1-m) x (ps+1-m) block
int coax_size(int pt, int ps);
int coax_offset(int pt, int ps, int m);

// Scaled spherical Bessel jhat_n(x) = j_n(x) (2n+1)!! / x^n and Hankel
// hhat_n(x) = h_n(x) x^{n+1} / (2n-1)!!, n = 0..p.
void jhat_host(double x, int p, double* out);
void hhat_host(double x, int p, double* re, double* im);

// Rotation about y: E^n(theta) with Y_n^m(R_y(theta) u) = sum_m' E[m][m'] Y_n^m'(u).
void rotation_table(double theta, int p, double* out /* rot_size(p) */);
// the rotation's blocks in the (u, v) basis of a degree block (tables.cpp)
void rotation_uv_blocks(const double* E, int n, bool transpose, std::vector<double>& Wu, std::vector<double>& Wv);

// Gaunt-type coaxial coupling integrals, computed once per plan for degree pmax.
struct Gaunt {
    int p = -1;
    std::vector<double> g;   // [m][n'][l][n], n,n' in [m,p], l in [0, 2p]
    double at(int m, int n2, int l, int n) const {
        const int P1 = p + 1, L1 = 2 * p + 1;
        return g[(((size_t)m * P1 + n2) * L1 + l) * P1 + n];
    }
};
Gaunt make_gaunt(int p);

// Scaled coaxial translation operators along +z (complex, interleaved re/im).
// kind 0: M2L (source radius Rs, target radius Rt, distance t)
// kind 1: M2M (child radius Rs -> parent radius Rt)
// kind 2: L2L (parent radius Rs -> child radius Rt)
void coax_table(int kind, const Gaunt& G, double k, double t, double Rs, double Rt, int ps, int pt,
                double* out /* 2*coax_size(pt,ps) */);

}  // namespace hfmm
