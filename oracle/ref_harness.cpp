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
//   time IN theta softening_m nsample [hashed [reps]]
//       IN : int64 n, float x[n], y[n], z[n], m[n]
//       Times the serial Tree::insert of all n bodies and the OpenMP
//       treeInteraction loop over every (n/nsample)-th body, reps times (a new
//       tree each time); prints one JSON line with the medians and every rep.
//       hashed = 1 walks the bodies in key order, the reference's `hashed` path
//       (Interaction.cpp:156-177,190-197): Body::discretizePosition +
//       Body::getKey per body into a std::map<key, Body*> (a repeated key keeps
//       the last body), and the sample is every (n/nsample)-th body of that
//       order.  The map is timed on its own (key_s), as the reference leaves it
//       outside its build and force timers.
//   stars IN OUT softening_m time_step
//       IN : int64 ns, int64 nb, then for the ns suns and the nb bodies in turn
//            double m[], x[], y[], z[], vx[], vy[], vz[]
//       OUT: int64 ns, nb; the velocity increments of the star loop of
//            Interaction::interactBodies (Interaction.cpp:111-124) for the suns
//            (3 ns doubles) and bodies (3 nb); then, after
//            Interaction::updateBodies (:219-255), sun positions and
//            velocities, body positions and velocities (3 doubles each).
#include <algorithm>
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

int cmd_time(const char *in, double theta, double soft, int64_t nsample, int hashed, int reps)
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
    if (nsample <= 0 || nsample > n) nsample = n;
    int64_t stride = n / nsample;
    if (stride < 1) stride = 1;
    const int64_t ns = (norder + stride - 1) / stride;
    if (reps < 1) reps = 1;
    std::vector<double> build_s(reps), force_s(reps);
    int64_t inserted = 0;
    for (int r = 0; r < reps; ++r) {  // one step each: a fresh tree, then the sampled walks
        auto t0 = std::chrono::steady_clock::now();
        Tree *tree = new Tree(Octant(cx, cy, cz, len));
        inserted = 0;
        for (int64_t i = 0; i < n; ++i)
            if (tree->getOctant().contains(bods[i].position)) { tree->insert(&bods[i]); ++inserted; }
        auto t1 = std::chrono::steady_clock::now();
        #pragma omp parallel for
        for (int64_t k = 0; k < ns; ++k) {
            Body *b = order[k * stride];
            if (tree->getOctant().contains(b->position)) inter.treeInteraction(tree, b);
        }
        auto t2 = std::chrono::steady_clock::now();
        build_s[r] = std::chrono::duration<double>(t1 - t0).count();
        force_s[r] = std::chrono::duration<double>(t2 - t1).count();
        delete tree;
    }
    auto median = [](std::vector<double> v) {
        std::sort(v.begin(), v.end());
        const size_t h = v.size() / 2;
        return v.size() % 2 ? v[h] : 0.5 * (v[h - 1] + v[h]);
    };
    auto list = [](const std::vector<double> &v) {
        std::string s = "[";
        char buf[32];
        for (size_t i = 0; i < v.size(); ++i) {
            snprintf(buf, sizeof buf, "%s%.6f", i ? ", " : "", v[i]);
            s += buf;
        }
        return s + "]";
    };
    printf("{\"n\": %lld, \"inserted\": %lld, \"build_s\": %.6f, \"force_sample_s\": %.6f, "
           "\"build_reps_s\": %s, \"force_reps_s\": %s, "
           "\"nsample\": %lld, \"threads\": %d, \"theta\": %g, \"order\": \"%s\", \"ordered\": %lld, "
           "\"key_s\": %.6f}\n",
           (long long)n, (long long)inserted, median(build_s), median(force_s), list(build_s).c_str(),
           list(force_s).c_str(), (long long)ns, omp_get_max_threads(), theta, hashed ? "key" : "index",
           (long long)norder, key_s);
    delete[] bods;
    return 0;
}

// Star ("sun") forces and the update of Interaction::interactBodies /
// Interaction::updateBodies.  The loop over sun pairs and sun-body pairs is the
// one of Interaction.cpp:111-124, run on one thread: the reference's OpenMP
// loops there all add into the same sun's acceleration, and one thread gives
// the order those additions are written in.  Every interaction is the
// reference's own Interaction::singleInteraction, the update its own
// Interaction::updateBodies.
int cmd_stars(const char *in, const char *out, double soft, int time_step)
{
    FILE *f = fopen(in, "rb");
    if (!f) return 2;
    int64_t ns = 0, nb = 0;
    if (fread(&ns, 8, 1, f) != 1 || fread(&nb, 8, 1, f) != 1) return 2;
    std::vector<double> s[7], b[7];  // m, x, y, z, vx, vy, vz
    for (auto &v : s)
        if (!read_vec(f, v, ns)) return 2;
    for (auto &v : b)
        if (!read_vec(f, v, nb)) return 2;
    fclose(f);
    NUM_BODIES = (int)nb; NUM_SUNS = (int)ns; TIME_STEP = time_step; SOFTENING = soft;
    Interaction inter(false);
    LOGCFG.level = ERROR;
    std::vector<Body> suns(ns), bods(nb);
    auto load = [](std::vector<Body> &to, std::vector<double> *v) {
        for (size_t i = 0; i < to.size(); ++i) {
            to[i].mass = v[0][i];
            to[i].position = Vector3D(v[1][i], v[2][i], v[3][i]);
            to[i].velocity = Vector3D(v[4][i], v[5][i], v[6][i]);
            to[i].acceleration = Vector3D(0.0, 0.0, 0.0);
        }
    };
    load(suns, s);
    load(bods, b);
    omp_set_num_threads(1);
    for (int si = 0; si < NUM_SUNS; si++) {
        for (int ss = 0; ss < NUM_SUNS; ss++)
            if (si != ss) inter.singleInteraction(&suns[si], &suns[ss], false, false);
        for (int bi = 0; bi < NUM_BODIES; bi++) inter.singleInteraction(&suns[si], &bods[bi], true);
    }
    FILE *g = fopen(out, "wb");
    if (!g) return 3;
    auto put3 = [g](const std::vector<Body> &v, int what) {
        for (const Body &q : v) {
            const Vector3D &w = what == 0 ? q.acceleration : what == 1 ? q.position : q.velocity;
            double a[3] = {w.x, w.y, w.z};
            fwrite(a, 8, 3, g);
        }
    };
    fwrite(&ns, 8, 1, g);
    fwrite(&nb, 8, 1, g);
    put3(suns, 0);
    put3(bods, 0);
    inter.updateBodies(suns.data(), bods.data());
    put3(suns, 1);
    put3(suns, 2);
    put3(bods, 1);
    put3(bods, 2);
    fclose(g);
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
    if (c == "time" && argc >= 6 && argc <= 8)
        return cmd_time(argv[2], atof(argv[3]), atof(argv[4]), atoll(argv[5]), argc >= 7 ? atoi(argv[6]) : 0,
                        argc == 8 ? atoi(argv[7]) : 1);
    if (c == "stars" && argc == 6) return cmd_stars(argv[2], argv[3], atof(argv[4]), atoi(argv[5]));
    fprintf(stderr, "bad arguments\n");
    return 1;
}
