// ref_harness.cpp -- driver around the reference's own OpenMP Barnes-Hut sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the unmodified sources under 3D/OpenMP/OpenMP_NBody/src (Body, Octant, Tree,
// Vector3D, Interaction, Logger, Utils, Timer) into oracle/_ref/ref_harness.
// Nothing from the reference is copied here: this file only calls it.
//
// The reference's main.cpp (which needs Boost) normally defines the program
// globals declared extern in Constants.h and Logger.h; this driver defines
// them instead and sets them per command.  Interaction::treeInteraction is a
// private member, so the reference headers are included with private members
// opened to this translation unit only.
//
// Commands (binary little-endian files):
//   keys IN OUT
//       IN : int64 n, float x[n], y[n], z[n]
//       OUT: int64 n, uint64 key[n], int64 k, int32 order[k]
//       Bounding box as Interaction.cpp:130-147 computes it (min/max initialised
//       from the first body), Body::discretizePosition + Body::getKey per body,
//       and the std::map<key, Body*> order of Interaction.cpp:156-177.
//   tree IN OUT theta softening_m cx cy cz len
//       IN : int64 n, double x[n], y[n], z[n], m[n]
//       OUT: int64 n, double dv[3n], int64 nodes, nodes x {int32 level, int32 external,
//            double cx, cy, cz, len, mass, comx, comy, comz} in depth-first UNW..LSE order
//       Root Octant(cx,cy,cz,len); Tree::insert of every contained body in index
//       order (Interaction.cpp:179-186) and Interaction::treeInteraction per body
//       (:203-208) with TIME_STEP = 1 and SOFTENING = softening_m.  dv is the
//       velocity increment the reference accumulates in Body::acceleration.
//   time IN theta softening_m nsample [hashed]
//       IN : int64 n, float x[n], y[n], z[n], m[n]
//       Times the serial Tree::insert of all n bodies and the OpenMP
//       treeInteraction loop over every (n/nsample)-th body; prints one JSON line.
//       hashed = 1 walks the bodies in key order, the reference's `hashed` path
//       (Interaction.cpp:156-177,190-197): Body::discretizePosition +
//       Body::getKey per body into a std::map<key, Body*> (a repeated key keeps
//       the last body), and the sample is every (n/nsample)-th body of that
//       order.  The map is timed on its own (key_s), as the reference leaves it
//       outside its build and force timers.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <utility>
#include <vector>
#include <omp.h>

#define private public
#include "Interaction.h"
#undef private

int NUM_BODIES, NUM_SUNS, TIME_STEP, STEP_COUNT, WIDTH, HEIGHT;
double SYSTEM_SIZE, SYSTEM_THICKNESS, INNER_BOUND, SOFTENING, SOLAR_MASS, EXTRA_MASS,
    MAX_DISTANCE, BINARY_SEPARATION;
bool BINARY;
structlog LOGCFG = {};

namespace {

template <class T> bool read_vec(FILE *f, std::vector<T> &v, size_t n)
{
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

struct Span { double x0, y0, z0, span; };

Span span_of(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &z)
{
    double xMin = x[0], xMax = x[0], yMin = y[0], yMax = y[0], zMin = z[0], zMax = z[0];
    for (size_t i = 0; i < x.size(); ++i) {
        xMin = x[i] < xMin ? x[i] : xMin; yMin = y[i] < yMin ? y[i] : yMin;
        zMin = z[i] < zMin ? z[i] : zMin; xMax = x[i] > xMax ? x[i] : xMax;
        yMax = y[i] > yMax ? y[i] : yMax; zMax = z[i] > zMax ? z[i] : zMax;
    }
    double xSpan = std::fabs(xMax - xMin), ySpan = std::fabs(yMax - yMin), zSpan = std::fabs(zMax - zMin);
    double maxSpan = xSpan > ySpan ? xSpan : ySpan;
    maxSpan = maxSpan < zSpan ? zSpan : maxSpan;
    return {xMin, yMin, zMin, maxSpan};
}

int cmd_keys(const char *in, const char *out)
{
    FILE *f = fopen(in, "rb");
    if (!f) return 2;
    int64_t n = 0;
    if (fread(&n, 8, 1, f) != 1) return 2;
    std::vector<float> xf, yf, zf;
    if (!read_vec(f, xf, n) || !read_vec(f, yf, n) || !read_vec(f, zf, n)) return 2;
    fclose(f);
    std::vector<double> x(xf.begin(), xf.end()), y(yf.begin(), yf.end()), z(zf.begin(), zf.end());
    Span s = span_of(x, y, z);
    std::vector<Body> bods(n);
    std::vector<uint64_t> keys(n);
    std::map<uint64_t, Body *> bodies;
    for (int64_t i = 0; i < n; ++i) {
        bods[i].position = Vector3D(x[i], y[i], z[i]);
        bods[i].discretizePosition(s.span, Vector3D(s.x0, s.y0, s.z0));
        keys[i] = bods[i].getKey();
        bodies[keys[i]] = &bods[i];
    }
    std::vector<int32_t> order;
    for (auto const &b : bodies) order.push_back((int32_t)(b.second - bods.data()));
    FILE *g = fopen(out, "wb");
    if (!g) return 3;
    int64_t k = (int64_t)order.size();
    fwrite(&n, 8, 1, g);
    fwrite(keys.data(), 8, n, g);
    fwrite(&k, 8, 1, g);
    fwrite(order.data(), 4, k, g);
    fclose(g);
    return 0;
}

struct NodeRec { int32_t level, external; double d[8]; };

void dump(Tree *t, int level, std::vector<NodeRec> &out)
{
    NodeRec r;
    r.level = level;
    r.external = t->isExternal() ? 1 : 0;
    r.d[0] = t->octant.center.x; r.d[1] = t->octant.center.y; r.d[2] = t->octant.center.z;
    r.d[3] = t->octant.getLength();
    r.d[4] = t->centerOfMass.mass;
    r.d[5] = t->centerOfMass.position.x; r.d[6] = t->centerOfMass.position.y;
    r.d[7] = t->centerOfMass.position.z;
    out.push_back(r);
    Tree *ch[8] = {t->UNW, t->UNE, t->USW, t->USE, t->LNW, t->LNE, t->LSW, t->LSE};
    for (int s = 0; s < 8; ++s)
        if (ch[s] != NULL) dump(ch[s], level + 1, out);
}

int cmd_tree(const char *in, const char *out, double theta, double soft, double cx, double cy,
             double cz, double len)
{
    FILE *f = fopen(in, "rb");
    if (!f) return 2;
    int64_t n = 0;
    if (fread(&n, 8, 1, f) != 1) return 2;
    std::vector<double> x, y, z, m;
    if (!read_vec(f, x, n) || !read_vec(f, y, n) || !read_vec(f, z, n) || !read_vec(f, m, n)) return 2;
    fclose(f);
    NUM_BODIES = (int)n; NUM_SUNS = 0; TIME_STEP = 1; SOFTENING = soft; MAX_DISTANCE = theta;
    Interaction inter(false);
    LOGCFG.level = ERROR;
    std::vector<Body> bods(n);
    for (int64_t i = 0; i < n; ++i) {
        bods[i].mass = m[i];
        bods[i].position = Vector3D(x[i], y[i], z[i]);
    }
    Tree *tree = new Tree(Octant(cx, cy, cz, len));
    for (int64_t i = 0; i < n; ++i)
        if (tree->getOctant().contains(bods[i].position)) tree->insert(&bods[i]);
    for (int64_t i = 0; i < n; ++i)
        if (tree->getOctant().contains(bods[i].position)) inter.treeInteraction(tree, &bods[i]);
    std::vector<NodeRec> nodes;
    dump(tree, 0, nodes);
    delete tree;
    FILE *g = fopen(out, "wb");
    if (!g) return 3;
    fwrite(&n, 8, 1, g);
    for (int64_t i = 0; i < n; ++i) {
        double a[3] = {bods[i].acceleration.x, bods[i].acceleration.y, bods[i].acceleration.z};
        fwrite(a, 8, 3, g);
    }
    int64_t nn = (int64_t)nodes.size();
    fwrite(&nn, 8, 1, g);
    for (auto const &r : nodes) {
        fwrite(&r.level, 4, 1, g);
        fwrite(&r.external, 4, 1, g);
        fwrite(r.d, 8, 8, g);
    }
    fclose(g);
    return 0;
}

int cmd_time(const char *in, double theta, double soft, int64_t nsample, int hashed)
{
    FILE *f = fopen(in, "rb");
    if (!f) return 2;
    int64_t n = 0;
    if (fread(&n, 8, 1, f) != 1) return 2;
    std::vector<float> xf, yf, zf, mf;
    if (!read_vec(f, xf, n) || !read_vec(f, yf, n) || !read_vec(f, zf, n) || !read_vec(f, mf, n)) return 2;
    fclose(f);
    NUM_BODIES = (int)n; NUM_SUNS = 0; TIME_STEP = 1; SOFTENING = soft; MAX_DISTANCE = theta;
    Interaction inter(false);
    LOGCFG.level = ERROR;
    std::vector<double> x(xf.begin(), xf.end()), y(yf.begin(), yf.end()), z(zf.begin(), zf.end());
    Span s = span_of(x, y, z);
    // Root cube centred on the bounding box (the reference centres it on the
    // origin from an uninitialised min/max scan, Interaction.cpp:130-149).
    double cx = s.x0 + 0.5 * s.span, cy = s.y0 + 0.5 * s.span, cz = s.z0 + 0.5 * s.span;
    double len = s.span * (1.0 + 1e-9);
    Body *bods = new Body[n];
    for (int64_t i = 0; i < n; ++i) {
        bods[i].mass = mf[i];
        bods[i].position = Vector3D(x[i], y[i], z[i]);
    }
    std::vector<Body *> order;  // traversal order
    double key_s = 0.0;
    if (hashed) {
        auto k0 = std::chrono::steady_clock::now();
        std::map<uint64_t, Body *> bodies;
        for (int64_t i = 0; i < n; ++i) {
            bods[i].discretizePosition(s.span, Vector3D(s.x0, s.y0, s.z0));
            bodies[bods[i].getKey()] = &bods[i];
        }
        order.reserve(bodies.size());
        for (auto const &b : bodies) order.push_back(b.second);
        key_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - k0).count();
    } else {
        order.reserve(n);
        for (int64_t i = 0; i < n; ++i) order.push_back(&bods[i]);
    }
    const int64_t norder = (int64_t)order.size();
    auto t0 = std::chrono::steady_clock::now();
    Tree *tree = new Tree(Octant(cx, cy, cz, len));
    int64_t inserted = 0;
    for (int64_t i = 0; i < n; ++i)
        if (tree->getOctant().contains(bods[i].position)) { tree->insert(&bods[i]); ++inserted; }
    auto t1 = std::chrono::steady_clock::now();
    if (nsample <= 0 || nsample > n) nsample = n;
    int64_t stride = n / nsample;
    if (stride < 1) stride = 1;
    int64_t ns = (norder + stride - 1) / stride;
    #pragma omp parallel for
    for (int64_t k = 0; k < ns; ++k) {
        Body *b = order[k * stride];
        if (tree->getOctant().contains(b->position)) inter.treeInteraction(tree, b);
    }
    auto t2 = std::chrono::steady_clock::now();
    double build_s = std::chrono::duration<double>(t1 - t0).count();
    double force_s = std::chrono::duration<double>(t2 - t1).count();
    printf("{\"n\": %lld, \"inserted\": %lld, \"build_s\": %.6f, \"force_sample_s\": %.6f, "
           "\"nsample\": %lld, \"threads\": %d, \"theta\": %g, \"order\": \"%s\", \"ordered\": %lld, "
           "\"key_s\": %.6f}\n",
           (long long)n, (long long)inserted, build_s, force_s, (long long)ns, omp_get_max_threads(), theta,
           hashed ? "key" : "index", (long long)norder, key_s);
    delete tree;
    delete[] bods;
    return 0;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: ref_harness keys|tree|time ...\n");
        return 1;
    }
    std::string c = argv[1];
    if (c == "keys" && argc == 4) return cmd_keys(argv[2], argv[3]);
    if (c == "tree" && argc == 10)
        return cmd_tree(argv[2], argv[3], atof(argv[4]), atof(argv[5]), atof(argv[6]), atof(argv[7]),
                        atof(argv[8]), atof(argv[9]));
    if (c == "time" && (argc == 6 || argc == 7))
        return cmd_time(argv[2], atof(argv[3]), atof(argv[4]), atoll(argv[5]), argc == 7 ? atoi(argv[6]) : 0);
    fprintf(stderr, "bad arguments\n");
    return 1;
}
