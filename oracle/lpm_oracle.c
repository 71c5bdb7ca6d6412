/*
 * lpm_oracle.c -- CPU oracle for the spherical barotropic-vorticity (BVE)
 * Biot-Savart hot path of pbosler/lpmKokkos.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, the smoke check and the CPU-baseline
 * leg of bench.py load it to produce expected values.  The product library (lpmkokkos_amd) never links, loads or calls
 * anything here, and nothing here runs on a GPU.
 *
 * Every function restates the reference's arithmetic in the reference's
 * evaluation order, compiled with -ffp-contract=off so that no fused
 * multiply-add changes a rounding.  Citations are reference file:line.
 *
 * Parity pinning (what checks each part; see tests/test_oracle_*.py):
 *   geometry      -- tests/LpmGeometryUnitTests.cpp:145-165 known answers
 *                    (triArea = pi/3, polygonArea = 2pi/3, midpoint, cross,
 *                    barycenter), plus closed forms.
 *   mesh          -- seed file values, closed-form counts
 *                    (src/LpmMeshSeed.cpp:381-391), sum of areas = 4 pi,
 *                    Euler characteristic, and the published Y_5^4 Poisson
 *                    L-infinity table (scripts/poissonOutput.py:6-30) that
 *                    depends on every position, area and mask.
 *   pair sums     -- the same Poisson table (stream function), brute force in
 *                    long double on small meshes, u.x = 0, solid-body
 *                    rotation closed form, rotation equivariance.
 *   RK4           -- KokkosBlas scal/update known answers
 *                    (tests/LpmKokkosKernelsTests.cpp:19-86), solid-body
 *                    rotation exact trajectory (tests/LpmBVETest.cpp:186-191),
 *                    the face-vorticity stage quirk, Rossby-Haurwitz phase.
 *   ICs / PRNG    -- published SplitMix64 / xoshiro256** output vectors,
 *                    closed-form RH values.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* src/LpmDefs.hpp:82 */
static const double PI = 3.1415926535897932384626433832795027975;
#define NULL_IND (-1)

/* ------------------------------------------------------------------ */
/* geometry: src/LpmGeometry.hpp:279-468, src/LpmUtilities.hpp:14-46   */
/* ------------------------------------------------------------------ */

static inline double g_dot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; /* LpmGeometry.hpp:279-282 */
}
static inline void g_cross(double* c, const double* a, const double* b) {
  c[0] = a[1] * b[2] - a[2] * b[1]; /* LpmGeometry.hpp:291-295 */
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double g_mag(const double* v) { return sqrt(g_dot(v, v)); }
static inline void g_normalize(double* v) { /* LpmGeometry.hpp:360-363 */
  const double s = 1.0 / g_mag(v);
  v[0] *= s;
  v[1] *= s;
  v[2] *= s;
}
static inline double g_distance(const double* a, const double* b) {
  double cp[3]; /* LpmGeometry.hpp:370-376 */
  g_cross(cp, a, b);
  const double dp = g_dot(a, b);
  return atan2(g_mag(cp), dp);
}
static inline void g_midpoint(double* v, const double* a, const double* b) {
  v[0] = 0.5 * (a[0] + b[0]); /* LpmGeometry.hpp:427-433 */
  v[1] = 0.5 * (a[1] + b[1]);
  v[2] = 0.5 * (a[2] + b[2]);
  g_normalize(v);
}
/* LpmGeometry.hpp:412-422: sum rows, scale by 1.0/n, normalise */
static inline void g_barycenter(double* v, const double (*cv)[3], int n) {
  v[0] = 0.0;
  v[1] = 0.0;
  v[2] = 0.0;
  for (int i = 0; i < n; ++i) {
    v[0] += cv[i][0];
    v[1] += cv[i][1];
    v[2] += cv[i][2];
  }
  const double s = 1.0 / n;
  v[0] *= s;
  v[1] *= s;
  v[2] *= s;
  g_normalize(v);
}
/* L'Huilier: LpmGeometry.hpp:442-453 */
static inline double g_tri_area(const double* a, const double* b, const double* c) {
  const double s1 = g_distance(a, b);
  const double s2 = g_distance(b, c);
  const double s3 = g_distance(c, a);
  const double half_perim = 0.5 * (s1 + s2 + s3);
  const double zz = tan(0.5 * half_perim) * tan(0.5 * (half_perim - s1)) *
                    tan(0.5 * (half_perim - s2)) * tan(0.5 * (half_perim - s3));
  return 4 * atan(sqrt(zz));
}
/* LpmGeometry.hpp:461-468 */
static inline double g_polygon_area(const double* ctr, const double (*verts)[3], int n) {
  double ar = 0;
  for (int i = 0; i < n; ++i) ar += g_tri_area(ctr, verts[i], verts[(i + 1) % n]);
  return ar;
}
/* src/LpmUtilities.hpp:14-46 */
static double g_atan4(const double y, const double x) {
  double result = 0.0;
  if (x == 0.0) {
    if (y > 0.0)
      result = 0.5 * PI;
    else if (y < 0.0)
      result = 1.5 * PI;
    else if (y == 0.0)
      result = 0.0;
  } else if (y == 0) {
    if (x > 0.0)
      result = 0.0;
    else if (x < 0.0)
      result = PI;
  } else {
    const double theta = atan2(fabs(y), fabs(x));
    if (x > 0.0 && y > 0.0)
      result = theta;
    else if (x < 0.0 && y > 0.0)
      result = PI - theta;
    else if (x < 0.0 && y < 0.0)
      result = PI + theta;
    else if (x > 0.0 && y < 0.0)
      result = 2.0 * PI - theta;
  }
  return result;
}
/* LpmGeometry.hpp:216-228 */
static inline double g_latitude(const double* v) { return atan2(v[2], sqrt(v[0] * v[0] + v[1] * v[1])); }
static inline double g_longitude(const double* v) { return g_atan4(v[1], v[0]); }

ORC_EXPORT double orc_tri_area(const double* a, const double* b, const double* c) { return g_tri_area(a, b, c); }
ORC_EXPORT double orc_polygon_area(const double* ctr, const double* verts, int n) {
  return g_polygon_area(ctr, (const double(*)[3])verts, n);
}
ORC_EXPORT void orc_midpoint(double* v, const double* a, const double* b) { g_midpoint(v, a, b); }
ORC_EXPORT void orc_barycenter(double* v, const double* cv, int n) { g_barycenter(v, (const double(*)[3])cv, n); }
ORC_EXPORT void orc_cross(double* c, const double* a, const double* b) { g_cross(c, a, b); }
ORC_EXPORT double orc_distance(const double* a, const double* b) { return g_distance(a, b); }
ORC_EXPORT double orc_latitude(const double* v) { return g_latitude(v); }
ORC_EXPORT double orc_longitude(const double* v) { return g_longitude(v); }

/* ------------------------------------------------------------------ */
/* icosahedral seed file: src/LpmMeshSeed.cpp:20-232                   */
/* ------------------------------------------------------------------ */
#define SEED_NV 12
#define SEED_NF 20
#define SEED_NE 30

/* Line-oriented reader for the seed format: a header line, 32 coordinate
 * lines, then "edgeO", "faceverts", "faceedges", "vertEdges" sections
 * (src/LpmMeshSeed.cpp:43-230).  Returns 0 on success. */
ORC_EXPORT int orc_seed_read(const char* path, double* crds /*32x3*/, int* edges /*30x4*/,
                             int* faceverts /*20x3*/, int* faceedges /*20x3*/) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[1024];
  int lineno = 0, edgeHdr = -1, fvHdr = -1, feHdr = -1;
  int nc = 0, ne = 0, nfv = 0, nfe = 0;
  const int ncrds = SEED_NV + SEED_NF;
  while (fgets(line, sizeof line, f)) {
    ++lineno;
    if (strstr(line, "edgeO")) edgeHdr = lineno;
    if (strstr(line, "faceverts")) fvHdr = lineno;
    if (strstr(line, "faceedges")) feHdr = lineno;
    if (lineno > 1 && lineno < ncrds + 2) {
      double x, y, z;
      if (sscanf(line, "%lf %lf %lf", &x, &y, &z) != 3) { fclose(f); return -2; }
      crds[3 * nc + 0] = x;
      crds[3 * nc + 1] = y;
      crds[3 * nc + 2] = z;
      ++nc;
    } else if (edgeHdr > 0 && lineno > edgeHdr && lineno < edgeHdr + SEED_NE + 1) {
      int o, d, l, r;
      if (sscanf(line, "%d %d %d %d", &o, &d, &l, &r) != 4) { fclose(f); return -3; }
      edges[4 * ne + 0] = o;
      edges[4 * ne + 1] = d;
      edges[4 * ne + 2] = l;
      edges[4 * ne + 3] = r;
      ++ne;
    } else if (fvHdr > 0 && lineno > fvHdr && lineno < fvHdr + SEED_NF + 1) {
      int a, b, c;
      if (sscanf(line, "%d %d %d", &a, &b, &c) != 3) { fclose(f); return -4; }
      faceverts[3 * nfv + 0] = a;
      faceverts[3 * nfv + 1] = b;
      faceverts[3 * nfv + 2] = c;
      ++nfv;
    } else if (feHdr > 0 && lineno > feHdr && lineno < feHdr + SEED_NF + 1) {
      int a, b, c;
      if (sscanf(line, "%d %d %d", &a, &b, &c) != 3) { fclose(f); return -5; }
      faceedges[3 * nfe + 0] = a;
      faceedges[3 * nfe + 1] = b;
      faceedges[3 * nfe + 2] = c;
      ++nfe;
    }
  }
  fclose(f);
  return (nc == ncrds && ne == SEED_NE && nfv == SEED_NF && nfe == SEED_NF) ? 0 : -6;
}

/* src/LpmMeshSeed.cpp:296-305 and :381-391 */
ORC_EXPORT void orc_icos_alloc(int lev, int* nv, int* ne, int* nf) {
  long p = 1;
  for (int i = 0; i < lev; ++i) p *= 4;
  *nv = (int)(2 + 10 * p);
  int faces = 0, edges = 0;
  long q = 1;
  for (int i = 0; i <= lev; ++i) {
    const int nvi = (int)(2 + 10 * q), nfi = (int)(20 * q);
    faces += nfi;
    edges += nvi + nfi - 2;
    q *= 4;
  }
  *nf = faces;
  *ne = edges;
}

/* ------------------------------------------------------------------ */
/* tree refinement: src/LpmPolyMesh2d.cpp:30-44,                       */
/* src/LpmFaces.cpp:115-248, src/LpmEdges.cpp:12-85                    */
/* ------------------------------------------------------------------ */
typedef struct {
  int nmaxv, nmaxe, nmaxf;
  int nv, ne, nf;
  double (*vx)[3], (*vlag)[3];
  int *eo, *ed, *el, *er, *ep, (*ek)[2];
  int (*fv)[3], (*fe)[3], *fctr, *fpar, (*fk)[4];
  double* farea;
  unsigned char* fmask;
  double (*fx)[3], (*flag)[3];
  int nfc; /* face coordinate count */
} orc_mesh;

static void edge_insert(orc_mesh* m, int o, int d, int l, int r, int prt) {
  const int ins = m->ne; /* src/LpmEdges.cpp:12-24 */
  m->eo[ins] = o;
  m->ed[ins] = d;
  m->el[ins] = l;
  m->er[ins] = r;
  m->ep[ins] = prt;
  m->ek[ins][0] = NULL_IND;
  m->ek[ins][1] = NULL_IND;
  m->ne += 1;
}

static void edge_divide(orc_mesh* m, int ind) { /* src/LpmEdges.cpp:55-85 */
  const int crd_ins_pt = m->nv;
  const int edge_ins_pt = m->ne;
  double midpt[3], lagmidpt[3];
  g_midpoint(midpt, m->vx[m->eo[ind]], m->vx[m->ed[ind]]);
  g_midpoint(lagmidpt, m->vlag[m->eo[ind]], m->vlag[m->ed[ind]]);
  memcpy(m->vx[m->nv], midpt, sizeof midpt);
  memcpy(m->vlag[m->nv], lagmidpt, sizeof lagmidpt);
  m->nv += 1;
  edge_insert(m, m->eo[ind], crd_ins_pt, m->el[ind], m->er[ind], ind);
  edge_insert(m, crd_ins_pt, m->ed[ind], m->el[ind], m->er[ind], ind);
  m->ek[ind][0] = edge_ins_pt;
  m->ek[ind][1] = edge_ins_pt + 1;
}

static void face_divide(orc_mesh* m, int faceInd) { /* src/LpmFaces.cpp:115-248 */
  int nfv[4][3], nfe[4][3], kids[4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) nfv[i][j] = nfe[i][j] = NULL_IND;
  const int face_insert_pt = m->nf;
  for (int i = 0; i < 4; ++i) kids[i] = face_insert_pt + i;
  for (int i = 0; i < 3; ++i) nfv[i][i] = m->fv[faceInd][i];
  for (int i = 0; i < 3; ++i) {
    const int parentEdge = m->fe[faceInd][i];
    int ek0, ek1;
    if (m->ek[parentEdge][0] > 0) { /* LpmEdges.hpp:226-228 */
      ek0 = m->ek[parentEdge][0];
      ek1 = m->ek[parentEdge][1];
    } else {
      ek0 = m->ne;
      ek1 = m->ne + 1;
      edge_divide(m, parentEdge);
    }
    if (faceInd == m->el[m->fe[faceInd][i]]) { /* LpmFaces.hpp:248-250 */
      nfe[i][i] = ek0;
      m->el[ek0] = kids[i];
      nfe[(i + 1) % 3][i] = ek1;
      m->el[ek1] = kids[(i + 1) % 3];
    } else {
      nfe[i][i] = ek1;
      m->er[ek1] = kids[i];
      nfe[(i + 1) % 3][i] = ek0;
      m->er[ek0] = kids[(i + 1) % 3];
    }
    const int c1dest = m->ed[ek0];
    if (i == 0) {
      nfv[0][1] = c1dest;
      nfv[1][0] = c1dest;
      nfv[3][2] = c1dest;
    } else if (i == 1) {
      nfv[1][2] = c1dest;
      nfv[2][1] = c1dest;
      nfv[3][0] = c1dest;
    } else {
      nfv[2][0] = c1dest;
      nfv[0][2] = c1dest;
      nfv[3][1] = c1dest;
    }
  }
  const int edge_ins_pt = m->ne;
  for (int i = 0; i < 3; ++i) nfe[3][i] = edge_ins_pt + i;
  nfe[0][1] = edge_ins_pt + 1;
  nfe[1][2] = edge_ins_pt + 2;
  nfe[2][0] = edge_ins_pt;
  edge_insert(m, nfv[2][1], nfv[2][0], kids[3], kids[2], NULL_IND);
  edge_insert(m, nfv[0][2], nfv[0][1], kids[3], kids[0], NULL_IND);
  edge_insert(m, nfv[1][0], nfv[1][2], kids[3], kids[1], NULL_IND);

  double ctr[4][3], lagctr[4][3], ar[4];
  for (int i = 0; i < 4; ++i) {
    double vc[3][3], vl[3][3];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        vc[j][k] = m->vx[nfv[i][j]][k];
        vl[j][k] = m->vlag[nfv[i][j]][k];
      }
    g_barycenter(ctr[i], (const double(*)[3])vc, 3);
    g_barycenter(lagctr[i], (const double(*)[3])vl, 3);
    ar[i] = g_polygon_area(ctr[i], (const double(*)[3])vc, 3);
  }
  const int crd_ins_pt = m->nfc;
  for (int i = 0; i < 4; ++i) {
    memcpy(m->fx[m->nfc], ctr[i], sizeof ctr[i]);
    memcpy(m->flag[m->nfc], lagctr[i], sizeof lagctr[i]);
    m->nfc += 1;
    const int ins = m->nf; /* src/LpmFaces.cpp:9-27 */
    for (int j = 0; j < 3; ++j) {
      m->fv[ins][j] = nfv[i][j];
      m->fe[ins][j] = nfe[i][j];
    }
    for (int j = 0; j < 4; ++j) m->fk[ins][j] = NULL_IND;
    m->fctr[ins] = crd_ins_pt + i;
    m->fpar[ins] = faceInd;
    m->farea[ins] = ar[i];
    m->fmask[ins] = 0;
    m->nf += 1;
  }
  for (int i = 0; i < 4; ++i) m->fk[faceInd][i] = kids[i];
  m->farea[faceInd] = 0.0;
  m->fmask[faceInd] = 1;
}

/*
 * Build the level-`lev` icosahedral tree mesh from seed arrays
 * (PolyMesh2d::treeInit, src/LpmPolyMesh2d.cpp:30-44).
 *
 * Outputs (caller allocated with orc_icos_alloc sizes):
 *   vx, vlag  [nv][3]     vertex physical / Lagrangian coordinates
 *   fx, flag  [nf][3]     face-centre physical / Lagrangian coordinates
 *   farea     [nf]        areas (0 for divided faces)
 *   fmask     [nf]        1 for divided faces
 *   fverts    [nf][3]     face vertex indices
 *   fparent   [nf]        parent face index (-1 for seed faces)
 *   counts    [4]         nv, ne, nf, nleaves actually built
 */
ORC_EXPORT int orc_icos_tree_init(int lev, const double* scrds, const int* sedges, const int* sfaceverts,
                                  const int* sfaceedges, double* vx, double* vlag, double* fx, double* flag,
                                  double* farea, unsigned char* fmask, int* fverts, int* fparent, int* counts) {
  orc_mesh m;
  memset(&m, 0, sizeof m);
  orc_icos_alloc(lev, &m.nmaxv, &m.nmaxe, &m.nmaxf);
  m.vx = (double(*)[3])vx;
  m.vlag = (double(*)[3])vlag;
  m.fx = (double(*)[3])fx;
  m.flag = (double(*)[3])flag;
  m.farea = farea;
  m.fmask = fmask;
  m.fv = (int(*)[3])fverts;
  m.fpar = fparent;
  m.eo = malloc(sizeof(int) * m.nmaxe);
  m.ed = malloc(sizeof(int) * m.nmaxe);
  m.el = malloc(sizeof(int) * m.nmaxe);
  m.er = malloc(sizeof(int) * m.nmaxe);
  m.ep = malloc(sizeof(int) * m.nmaxe);
  m.ek = malloc(sizeof(int) * 2 * m.nmaxe);
  m.fe = malloc(sizeof(int) * 3 * m.nmaxf);
  m.fctr = malloc(sizeof(int) * m.nmaxf);
  m.fk = malloc(sizeof(int) * 4 * m.nmaxf);
  /* seedInit: src/LpmPolyMesh2d.cpp:10-18 */
  for (int i = 0; i < SEED_NV; ++i)
    for (int j = 0; j < 3; ++j) m.vx[i][j] = m.vlag[i][j] = scrds[3 * i + j];
  m.nv = SEED_NV;
  for (int i = 0; i < SEED_NE; ++i) {
    m.eo[i] = sedges[4 * i + 0];
    m.ed[i] = sedges[4 * i + 1];
    m.el[i] = sedges[4 * i + 2];
    m.er[i] = sedges[4 * i + 3];
    m.ep[i] = NULL_IND;
    m.ek[i][0] = m.ek[i][1] = NULL_IND;
  }
  m.ne = SEED_NE;
  for (int i = 0; i < SEED_NF; ++i) { /* src/LpmFaces.cpp:43-63 */
    for (int j = 0; j < 3; ++j) {
      m.fv[i][j] = sfaceverts[3 * i + j];
      m.fe[i][j] = sfaceedges[3 * i + j];
    }
    m.fctr[i] = i;
    /* MeshSeed::faceArea, src/LpmMeshSeed.cpp:307-320 */
    double vc[3][3];
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) vc[k][j] = scrds[3 * sfaceverts[3 * i + k] + j];
    m.farea[i] = g_polygon_area(&scrds[3 * (SEED_NV + i)], (const double(*)[3])vc, 3);
    m.fpar[i] = NULL_IND;
    for (int j = 0; j < 4; ++j) m.fk[i][j] = NULL_IND;
    m.fmask[i] = 0;
    for (int j = 0; j < 3; ++j) m.fx[i][j] = m.flag[i][j] = scrds[3 * (SEED_NV + i) + j];
  }
  m.nf = SEED_NF;
  m.nfc = SEED_NF;
  for (int l = 0; l < lev; ++l) {
    const int stop = m.nf;
    for (int j = 0; j < stop; ++j)
      if (!(m.fk[j][0] > 0)) face_divide(&m, j);
  }
  int nleaves = 0;
  for (int i = 0; i < m.nf; ++i) nleaves += !m.fmask[i];
  counts[0] = m.nv;
  counts[1] = m.ne;
  counts[2] = m.nf;
  counts[3] = nleaves;
  free(m.eo); free(m.ed); free(m.el); free(m.er); free(m.ep); free(m.ek);
  free(m.fe); free(m.fctr); free(m.fk);
  return 0;
}

/* ------------------------------------------------------------------ */
/* cubed-sphere seed and quadrilateral tree:                           */
/* MeshSeed<CubedSphereSeed> (src/LpmMeshSeed.hpp:111-126,             */
/* src/LpmMeshSeed.cpp:20-232, 296-320, 369-379) and                   */
/* FaceDivider<SphereGeometry,QuadFace>::divide (src/LpmFaces.cpp:251-381)*/
/* ------------------------------------------------------------------ */
#define CUBE_NV 8
#define CUBE_NF 6
#define CUBE_NE 12

/* The seed format with quadrilateral faces: a header line, 14 coordinate
 * lines (8 vertices, then 6 face centres), then the "edgeO" (12 lines),
 * "faceverts" (6 x 4) and "faceedges" (6 x 4) sections.  0 on success. */
ORC_EXPORT int orc_cubed_seed_read(const char* path, double* crds /*14x3*/, int* edges /*12x4*/,
                                   int* faceverts /*6x4*/, int* faceedges /*6x4*/) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[1024];
  int lineno = 0, edgeHdr = -1, fvHdr = -1, feHdr = -1;
  int nc = 0, ne = 0, nfv = 0, nfe = 0;
  const int ncrds = CUBE_NV + CUBE_NF;
  while (fgets(line, sizeof line, f)) {
    ++lineno;
    if (strstr(line, "edgeO")) edgeHdr = lineno;
    if (strstr(line, "faceverts")) fvHdr = lineno;
    if (strstr(line, "faceedges")) feHdr = lineno;
    if (lineno > 1 && lineno < ncrds + 2) {
      if (sscanf(line, "%lf %lf %lf", &crds[3 * nc], &crds[3 * nc + 1], &crds[3 * nc + 2]) != 3) { fclose(f); return -2; }
      ++nc;
    } else if (edgeHdr > 0 && lineno > edgeHdr && lineno < edgeHdr + CUBE_NE + 1) {
      if (sscanf(line, "%d %d %d %d", &edges[4 * ne], &edges[4 * ne + 1], &edges[4 * ne + 2], &edges[4 * ne + 3]) != 4) {
        fclose(f);
        return -3;
      }
      ++ne;
    } else if (fvHdr > 0 && lineno > fvHdr && lineno < fvHdr + CUBE_NF + 1) {
      int* v = &faceverts[4 * nfv];
      if (sscanf(line, "%d %d %d %d", &v[0], &v[1], &v[2], &v[3]) != 4) { fclose(f); return -4; }
      ++nfv;
    } else if (feHdr > 0 && lineno > feHdr && lineno < feHdr + CUBE_NF + 1) {
      int* e = &faceedges[4 * nfe];
      if (sscanf(line, "%d %d %d %d", &e[0], &e[1], &e[2], &e[3]) != 4) { fclose(f); return -5; }
      ++nfe;
    }
  }
  fclose(f);
  return (nc == ncrds && ne == CUBE_NE && nfv == CUBE_NF && nfe == CUBE_NF) ? 0 : -6;
}

/* src/LpmMeshSeed.cpp:296-305 with the counts of :369-379:
 * nverts(l) = 2 + 6 4^l, nfaces(l) = 6 4^l, nedges(l) = nverts + nfaces - 2. */
ORC_EXPORT void orc_cubed_alloc(int lev, int* nv, int* ne, int* nf) {
  long p = 1;
  for (int i = 0; i < lev; ++i) p *= 4;
  *nv = (int)(2 + 6 * p);
  int faces = 0, edges = 0;
  long q = 1;
  for (int i = 0; i <= lev; ++i) {
    const int nvi = (int)(2 + 6 * q), nfi = (int)(6 * q);
    faces += nfi;
    edges += nfi + nvi - 2;
    q *= 4;
  }
  *nf = faces;
  *ne = edges;
}

typedef struct {
  orc_mesh e;     /* vertices, edges, face coordinates, areas, masks, parents, kids */
  int (*fv4)[4];  /* face vertices */
  int (*fe4)[4];  /* face edges */
} orc_qmesh;

static void face_insert4(orc_qmesh* q, int ctr_ind, const int* verts, const int* edges, int prt, double area) {
  orc_mesh* m = &q->e; /* src/LpmFaces.cpp:9-27 */
  const int ins = m->nf;
  for (int j = 0; j < 4; ++j) {
    q->fv4[ins][j] = verts[j];
    q->fe4[ins][j] = edges[j];
    m->fk[ins][j] = NULL_IND;
  }
  m->fctr[ins] = ctr_ind;
  m->fpar[ins] = prt;
  m->farea[ins] = area;
  m->fmask[ins] = 0;
  m->nf += 1;
}

static void quad_face_divide(orc_qmesh* q, int faceInd) { /* src/LpmFaces.cpp:251-381 */
  orc_mesh* m = &q->e;
  int nfv[4][4], nfe[4][4], kids[4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) nfv[i][j] = nfe[i][j] = NULL_IND;
  const int face_insert_pt = m->nf;
  for (int i = 0; i < 4; ++i) kids[i] = face_insert_pt + i;
  for (int i = 0; i < 4; ++i) nfv[i][i] = q->fv4[faceInd][i];
  for (int i = 0; i < 4; ++i) {
    const int parentEdge = q->fe4[faceInd][i];
    int ek0, ek1;
    if (m->ek[parentEdge][0] > 0) {
      ek0 = m->ek[parentEdge][0];
      ek1 = m->ek[parentEdge][1];
    } else {
      ek0 = m->ne;
      ek1 = m->ne + 1;
      edge_divide(m, parentEdge);
    }
    if (faceInd == m->el[parentEdge]) {
      nfe[i][i] = ek0;
      m->el[ek0] = kids[i];
      nfe[(i + 1) % 4][i] = ek1;
      m->el[ek1] = kids[(i + 1) % 4];
    } else {
      nfe[i][i] = ek1;
      m->er[ek1] = kids[i];
      nfe[(i + 1) % 4][i] = ek0;
      m->er[ek0] = kids[(i + 1) % 4];
    }
    const int c1dest = m->ed[ek0];
    nfv[i][(i + 1) % 4] = c1dest;
    nfv[(i + 1) % 4][i] = c1dest;
  }
  /* the parent's centre becomes the children's shared vertex */
  const int pc = m->fctr[faceInd];
  const int bndry_insert_pt = m->nv;
  memcpy(m->vx[m->nv], m->fx[pc], sizeof(double) * 3);
  memcpy(m->vlag[m->nv], m->flag[pc], sizeof(double) * 3);
  m->nv += 1;
  for (int i = 0; i < 4; ++i) nfv[i][(i + 2) % 4] = bndry_insert_pt;
  const int edge_ins_pt = m->ne;
  edge_insert(m, nfv[0][1], nfv[0][2], kids[0], kids[1], NULL_IND);
  nfe[0][1] = edge_ins_pt;
  nfe[1][3] = edge_ins_pt;
  edge_insert(m, nfv[2][0], nfv[2][3], kids[3], kids[2], NULL_IND);
  nfe[2][3] = edge_ins_pt + 1;
  nfe[3][1] = edge_ins_pt + 1;
  edge_insert(m, nfv[2][1], nfv[2][0], kids[1], kids[2], NULL_IND);
  nfe[1][2] = edge_ins_pt + 2;
  nfe[2][0] = edge_ins_pt + 2;
  edge_insert(m, nfv[3][1], nfv[3][0], kids[0], kids[3], NULL_IND);
  nfe[0][2] = edge_ins_pt + 3;
  nfe[3][0] = edge_ins_pt + 3;
  double ctr[4][3], lagctr[4][3], ar[4];
  for (int i = 0; i < 4; ++i) {
    double vc[4][3], vl[4][3];
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 3; ++k) {
        vc[j][k] = m->vx[nfv[i][j]][k];
        vl[j][k] = m->vlag[nfv[i][j]][k];
      }
    g_barycenter(ctr[i], (const double(*)[3])vc, 4);
    g_barycenter(lagctr[i], (const double(*)[3])vl, 4);
    ar[i] = g_polygon_area(ctr[i], (const double(*)[3])vc, 4);
  }
  const int crd_ins_pt = m->nfc;
  for (int i = 0; i < 4; ++i) {
    memcpy(m->fx[m->nfc], ctr[i], sizeof ctr[i]);
    memcpy(m->flag[m->nfc], lagctr[i], sizeof lagctr[i]);
    m->nfc += 1;
    face_insert4(q, crd_ins_pt + i, nfv[i], nfe[i], faceInd, ar[i]);
  }
  for (int i = 0; i < 4; ++i) m->fk[faceInd][i] = kids[i];
  m->farea[faceInd] = 0.0;
  m->fmask[faceInd] = 1;
}

/* PolyMesh2d<CubedSphereSeed>::treeInit (src/LpmPolyMesh2d.cpp:30-44): the
 * outputs of orc_icos_tree_init with fverts [nf][4], sized by orc_cubed_alloc. */
ORC_EXPORT int orc_cubed_tree_init(int lev, const double* scrds, const int* sedges, const int* sfaceverts,
                                   const int* sfaceedges, double* vx, double* vlag, double* fx, double* flag,
                                   double* farea, unsigned char* fmask, int* fverts, int* fparent, int* counts) {
  orc_qmesh q;
  memset(&q, 0, sizeof q);
  orc_mesh* m = &q.e;
  orc_cubed_alloc(lev, &m->nmaxv, &m->nmaxe, &m->nmaxf);
  m->vx = (double(*)[3])vx;
  m->vlag = (double(*)[3])vlag;
  m->fx = (double(*)[3])fx;
  m->flag = (double(*)[3])flag;
  m->farea = farea;
  m->fmask = fmask;
  m->fpar = fparent;
  q.fv4 = (int(*)[4])fverts;
  m->eo = malloc(sizeof(int) * m->nmaxe);
  m->ed = malloc(sizeof(int) * m->nmaxe);
  m->el = malloc(sizeof(int) * m->nmaxe);
  m->er = malloc(sizeof(int) * m->nmaxe);
  m->ep = malloc(sizeof(int) * m->nmaxe);
  m->ek = malloc(sizeof(int) * 2 * m->nmaxe);
  q.fe4 = malloc(sizeof(int) * 4 * m->nmaxf);
  m->fctr = malloc(sizeof(int) * m->nmaxf);
  m->fk = malloc(sizeof(int) * 4 * m->nmaxf);
  for (int i = 0; i < CUBE_NV; ++i)
    for (int j = 0; j < 3; ++j) m->vx[i][j] = m->vlag[i][j] = scrds[3 * i + j];
  m->nv = CUBE_NV;
  for (int i = 0; i < CUBE_NE; ++i) {
    m->eo[i] = sedges[4 * i + 0];
    m->ed[i] = sedges[4 * i + 1];
    m->el[i] = sedges[4 * i + 2];
    m->er[i] = sedges[4 * i + 3];
    m->ep[i] = NULL_IND;
    m->ek[i][0] = m->ek[i][1] = NULL_IND;
  }
  m->ne = CUBE_NE;
  for (int i = 0; i < CUBE_NF; ++i) { /* src/LpmFaces.cpp:43-63 */
    double vc[4][3];
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 3; ++j) vc[k][j] = scrds[3 * sfaceverts[4 * i + k] + j];
    /* MeshSeed::faceArea, src/LpmMeshSeed.cpp:307-320 */
    face_insert4(&q, i, &sfaceverts[4 * i], &sfaceedges[4 * i], NULL_IND,
                 g_polygon_area(&scrds[3 * (CUBE_NV + i)], (const double(*)[3])vc, 4));
    for (int j = 0; j < 3; ++j) m->fx[i][j] = m->flag[i][j] = scrds[3 * (CUBE_NV + i) + j];
  }
  m->nfc = CUBE_NF;
  for (int l = 0; l < lev; ++l) {
    const int stop = m->nf;
    for (int j = 0; j < stop; ++j)
      if (!(m->fk[j][0] > 0)) quad_face_divide(&q, j);
  }
  int nleaves = 0;
  for (int i = 0; i < m->nf; ++i) nleaves += !m->fmask[i];
  counts[0] = m->nv;
  counts[1] = m->ne;
  counts[2] = m->nf;
  counts[3] = nleaves;
  free(m->eo); free(m->ed); free(m->el); free(m->er); free(m->ep); free(m->ek);
  free(q.fe4); free(m->fctr); free(m->fk);
  return 0;
}

/* ------------------------------------------------------------------ */
/* pair sums: src/LpmBVEKernels.hpp:21-305                             */
/* ------------------------------------------------------------------ */

/* biotSavart, src/LpmBVEKernels.hpp:36-44 */
static inline void bs_pair(double* u, const double* tgt_x, const double* src_x, double src_vort, double src_area) {
  g_cross(u, tgt_x, src_x);
  const double strength = -src_vort * src_area / (4 * PI * (1 - g_dot(src_x, tgt_x)));
  for (int j = 0; j < 3; ++j) u[j] *= strength;
}
/* greensFn, src/LpmBVEKernels.hpp:21-26 */
static inline double gf_pair(const double* tgt_x, const double* src_x, double src_vort, double src_area) {
  const double circ = -src_vort * src_area;
  return log(1 - g_dot(tgt_x, src_x)) * circ / (4 * PI);
}

/*
 * Velocity at targets tx[nt] from sources fx[nf] (BVEVertexVelocity,
 * BVEFaceVelocity: src/LpmBVEKernels.hpp:163-186, 285-305).  Each target sums
 * its sources sequentially in ascending j, starting from zero, adding an exact
 * zero for masked j (src/LpmKokkosUtil.hpp:25-28).  collocated != 0 skips
 * j == i (VelocityReduceCollocated, :212-234) and then tx must be fx.
 * `stride` > 1 evaluates only targets i = 0, stride, 2*stride, ... (the CPU
 * baseline's bounded sample); other rows of u are left untouched.
 */
ORC_EXPORT void orc_velocity(double* u, const double* tx, int nt, const double* fx, const double* fzeta,
                             const double* farea, const unsigned char* fmask, int nf, int collocated, int stride) {
  if (stride < 1) stride = 1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i += stride) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < nf; ++j) {
      double ut[3] = {0.0, 0.0, 0.0};
      if (!fmask[j] && !(collocated && i == j)) bs_pair(ut, &tx[3 * i], &fx[3 * j], fzeta[j], farea[j]);
      acc[0] += ut[0];
      acc[1] += ut[1];
      acc[2] += ut[2];
    }
    u[3 * i + 0] = acc[0];
    u[3 * i + 1] = acc[1];
    u[3 * i + 2] = acc[2];
  }
}

/* Stream function (BVEVertexStreamFn / BVEFaceStreamFn, :140-161, 263-283). */
ORC_EXPORT void orc_streamfn(double* psi, const double* tx, int nt, const double* fx, const double* fzeta,
                             const double* farea, const unsigned char* fmask, int nf, int collocated) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; ++i) {
    double acc = 0.0;
    for (int j = 0; j < nf; ++j) {
      double p = 0;
      if (!fmask[j] && !(collocated && i == j)) p = gf_pair(&tx[3 * i], &fx[3 * j], fzeta[j], farea[j]);
      acc += p;
    }
    psi[i] = acc;
  }
}

/*
 * Projected-sphere velocity used by the Lagrangian scheme: the same kernel
 * with 1 - x.y replaced by |y - x|^2 / 2, which equals it for points on the
 * unit sphere and has no cancellation.  Numerator cross(x, y).
 */
static void velocity_chord(double* u, const double* tx, int nt, const double* fx, const double* fzeta,
                           const double* farea, const unsigned char* fmask, int nf, int collocated, int stride) {
  if (stride < 1) stride = 1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i += stride) {
    double acc[3] = {0.0, 0.0, 0.0};
    const double* x = &tx[3 * i];
    for (int j = 0; j < nf; ++j) {
      if (fmask[j] || (collocated && i == j)) continue;
      const double* y = &fx[3 * j];
      const double d0 = y[0] - x[0], d1 = y[1] - x[1], d2 = y[2] - x[2];
      const double half_r2 = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2);
      double c[3];
      g_cross(c, x, y);
      const double strength = -fzeta[j] * farea[j] / (4 * PI * half_r2);
      acc[0] += c[0] * strength;
      acc[1] += c[1] * strength;
      acc[2] += c[2] * strength;
    }
    u[3 * i + 0] = acc[0];
    u[3 * i + 1] = acc[1];
    u[3 * i + 2] = acc[2];
  }
}
static void streamfn_chord(double* psi, const double* tx, int nt, const double* fx, const double* fzeta,
                           const double* farea, const unsigned char* fmask, int nf, int collocated) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; ++i) {
    double acc = 0.0;
    const double* x = &tx[3 * i];
    for (int j = 0; j < nf; ++j) {
      if (fmask[j] || (collocated && i == j)) continue;
      const double* y = &fx[3 * j];
      const double d0 = y[0] - x[0], d1 = y[1] - x[1], d2 = y[2] - x[2];
      const double half_r2 = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2);
      acc += log(half_r2) * (-fzeta[j] * farea[j]) / (4 * PI);
    }
    psi[i] = acc;
  }
}

ORC_EXPORT void orc_velocity_chord(double* u, const double* tx, int nt, const double* fx, const double* fzeta,
                                   const double* farea, const unsigned char* fmask, int nf, int collocated,
                                   int stride) {
  velocity_chord(u, tx, nt, fx, fzeta, farea, fmask, nf, collocated, stride);
}
ORC_EXPORT void orc_streamfn_chord(double* psi, const double* tx, int nt, const double* fx, const double* fzeta,
                                   const double* farea, const unsigned char* fmask, int nf, int collocated) {
  streamfn_chord(psi, tx, nt, fx, fzeta, farea, fmask, nf, collocated);
}

/* Brute-force long-double velocity/stream function (pins the pair sums on
 * small meshes against rounding in the double sums). */
ORC_EXPORT void orc_velocity_ld(double* u, double* psi, const double* tx, int nt, const double* fx,
                                const double* fzeta, const double* farea, const unsigned char* fmask, int nf,
                                int collocated) {
  const long double LPI = 3.1415926535897932384626433832795028841L;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; ++i) {
    long double a0 = 0, a1 = 0, a2 = 0, p = 0;
    const double* x = &tx[3 * i];
    for (int j = 0; j < nf; ++j) {
      if (fmask[j] || (collocated && i == j)) continue;
      const double* y = &fx[3 * j];
      const long double dot = (long double)x[0] * y[0] + (long double)x[1] * y[1] + (long double)x[2] * y[2];
      const long double c0 = (long double)x[1] * y[2] - (long double)x[2] * y[1];
      const long double c1 = (long double)x[2] * y[0] - (long double)x[0] * y[2];
      const long double c2 = (long double)x[0] * y[1] - (long double)x[1] * y[0];
      const long double s = -(long double)fzeta[j] * farea[j] / (4 * LPI * (1 - dot));
      a0 += c0 * s;
      a1 += c1 * s;
      a2 += c2 * s;
      p += logl(1 - dot) * (-(long double)fzeta[j] * farea[j]) / (4 * LPI);
    }
    u[3 * i + 0] = (double)a0;
    u[3 * i + 1] = (double)a1;
    u[3 * i + 2] = (double)a2;
    psi[i] = (double)p;
  }
}

/* ------------------------------------------------------------------ */
/* RK4: src/LpmBVERK4_Impl.hpp:11-130, src/LpmBVEKernels.hpp:307-320  */
/* ------------------------------------------------------------------ */

/* BVEVorticityTendency, src/LpmBVEKernels.hpp:317-319 */
static void tendency(double* dz, const double* vel, int n, double dt, double Omega) {
  for (int i = 0; i < n; ++i) dz[i] = -2.0 * Omega * vel[3 * i + 2] * dt;
}
/* KokkosBlas::scal(R, a, X): R = a*X */
static void scal(double* r, double a, const double* x, int n) {
  for (int i = 0; i < n; ++i) r[i] = a * x[i];
}
/* KokkosBlas::update(1, X, b, Y, 0, Z): Z = X + b*Y */
static void update(double* z, const double* x, double b, const double* y, int n) {
  for (int i = 0; i < n; ++i) z[i] = x[i] + b * y[i];
}
/* BVERK4Update, src/LpmBVERK4_Impl.hpp:11-39 */
static const double sixth = 1.0 / 6.0;
static const double third = 1.0 / 3.0;
static void rk4_update(double* x, const double* x1, const double* x2, const double* x3, const double* x4,
                       double* v, const double* v1, const double* v2, const double* v3, const double* v4, int n) {
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < 3; ++j)
      x[3 * i + j] += sixth * (x1[3 * i + j] + x4[3 * i + j]) + third * (x2[3 * i + j] + x3[3 * i + j]);
    v[i] += sixth * (v1[i] + v4[i]) + third * (v2[i] + v3[i]);
  }
}

static void vel_all(int scheme, double* vu, const double* vx, int nv, double* fu, const double* fx,
                    const double* fz, const double* fa, const unsigned char* fm, int nf) {
  if (scheme == 0) {
    orc_velocity(vu, vx, nv, fx, fz, fa, fm, nf, 0, 1);
    orc_velocity(fu, fx, nf, fx, fz, fa, fm, nf, 1, 1);
  } else {
    velocity_chord(vu, vx, nv, fx, fz, fa, fm, nf, 0, 1);
    velocity_chord(fu, fx, nf, fx, fz, fa, fm, nf, 1, 1);
  }
}

static void normalize_all(double* x, int n) {
  for (int i = 0; i < n; ++i) g_normalize(&x[3 * i]);
}
static void lagr_vort(double* zeta, const double* absv, const double* x, int n, double Omega) {
  for (int i = 0; i < n; ++i) zeta[i] = absv[i] - 2.0 * Omega * x[3 * i + 2];
}

/*
 * One BVERK4::advance_timestep (src/LpmBVERK4_Impl.hpp:42-130).
 * scheme 0 ("ref"): the reference step exactly -- stage 1 reuses the entry
 *   velocity, vorticity integrated with dzeta = -2 Omega w dt, the face
 *   vorticity update combines stages (1,2,4,4) as at :120-121, no
 *   renormalisation, and the end-of-step velocity solve.
 * scheme 1 ("lagr"): same stage velocities but every stage position and the
 *   updated position are normalised to the unit sphere, vorticity follows
 *   zeta = absv - 2 Omega z with absv = zeta0 + 2 Omega z0 conserved, and the
 *   kernel uses |y-x|^2/2 for 1 - x.y.  absv_v / absv_f are required then.
 * Arrays are AoS [n][3] for positions and velocities; nf counts tree faces.
 */
ORC_EXPORT void orc_rk4_step(int scheme, double dt, double Omega, int nv, double* vx, double* vzeta, double* vvel,
                             int nf, double* fx, double* fzeta, double* fvel, const double* fa,
                             const unsigned char* fm, const double* absv_v, const double* absv_f) {
  const size_t v3 = sizeof(double) * 3 * (size_t)nv, f3 = sizeof(double) * 3 * (size_t)nf;
  const size_t v1 = sizeof(double) * (size_t)nv, f1 = sizeof(double) * (size_t)nf;
  double *vk[4], *fk[4], *vzk[4], *fzk[4];
  for (int s = 0; s < 4; ++s) {
    vk[s] = malloc(v3);
    fk[s] = malloc(f3);
    vzk[s] = malloc(v1);
    fzk[s] = malloc(f1);
  }
  double* vxw = malloc(v3);
  double* fxw = malloc(f3);
  double* vzw = malloc(v1);
  double* fzw = malloc(f1);
  const double c[4] = {0.0, 0.5, 0.5, 1.0};
  for (int s = 0; s < 4; ++s) {
    if (s > 0) {
      update(vxw, vx, c[s], vk[s - 1], 3 * nv);
      update(fxw, fx, c[s], fk[s - 1], 3 * nf);
      if (scheme == 0) {
        update(vzw, vzeta, c[s], vzk[s - 1], nv);
        update(fzw, fzeta, c[s], fzk[s - 1], nf);
      } else {
        normalize_all(vxw, nv);
        normalize_all(fxw, nf);
        lagr_vort(vzw, absv_v, vxw, nv, Omega);
        lagr_vort(fzw, absv_f, fxw, nf, Omega);
      }
      vel_all(scheme, vvel, vxw, nv, fvel, fxw, fzw, fa, fm, nf);
    }
    scal(vk[s], dt, vvel, 3 * nv);
    scal(fk[s], dt, fvel, 3 * nf);
    tendency(vzk[s], vvel, nv, dt, Omega);
    tendency(fzk[s], fvel, nf, dt, Omega);
  }
  if (scheme == 0) {
    rk4_update(vx, vk[0], vk[1], vk[2], vk[3], vzeta, vzk[0], vzk[1], vzk[2], vzk[3], nv);
    rk4_update(fx, fk[0], fk[1], fk[2], fk[3], fzeta, fzk[0], fzk[1], fzk[3], fzk[3], nf);
  } else {
    rk4_update(vx, vk[0], vk[1], vk[2], vk[3], vzeta, vzk[0], vzk[1], vzk[2], vzk[3], nv);
    rk4_update(fx, fk[0], fk[1], fk[2], fk[3], fzeta, fzk[0], fzk[1], fzk[2], fzk[3], nf);
    normalize_all(vx, nv);
    normalize_all(fx, nf);
    lagr_vort(vzeta, absv_v, vx, nv, Omega);
    lagr_vort(fzeta, absv_f, fx, nf, Omega);
  }
  vel_all(scheme, vvel, vx, nv, fvel, fx, fzeta, fa, fm, nf);
  for (int s = 0; s < 4; ++s) {
    free(vk[s]); free(fk[s]); free(vzk[s]); free(fzk[s]);
  }
  free(vxw); free(fxw); free(vzw); free(fzw);
}

/* Exposed pieces of the step for the KokkosBlas known-answer tests. */
ORC_EXPORT void orc_blas_update(double* z, const double* x, double b, const double* y, int n) { update(z, x, b, y, n); }
ORC_EXPORT void orc_blas_scal(double* r, double a, const double* x, int n) { scal(r, a, x, n); }
ORC_EXPORT void orc_rk4_combine(double* x, const double* x1, const double* x2, const double* x3, const double* x4,
                                double* v, const double* v1, const double* v2, const double* v3, const double* v4,
                                int n) {
  rk4_update(x, x1, x2, x3, x4, v, v1, v2, v3, v4, n);
}

/* ------------------------------------------------------------------ */
/* initial conditions                                                  */
/* ------------------------------------------------------------------ */

/* SplitMix64 (Steele, Lea, Flood 2014) seeding xoshiro256** (Blackman and
 * Vigna 2018), as published. */
typedef struct { uint64_t s[4]; } orc_rng;
static uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t xoshiro_next(orc_rng* r) {
  uint64_t* s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}
static void rng_seed(orc_rng* r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}
/* uniform in [0,1): top 53 bits */
static double rng_uniform(orc_rng* r) { return (double)(xoshiro_next(r) >> 11) * 0x1.0p-53; }
/* standard normal: Box-Muller, cosine branch, one output per two uniforms */
static double rng_normal(orc_rng* r) {
  const double u1 = 1.0 - rng_uniform(r); /* (0,1] */
  const double u2 = rng_uniform(r);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
}

ORC_EXPORT uint64_t orc_splitmix64(uint64_t* state) { return splitmix64(state); }
ORC_EXPORT void orc_xoshiro_raw(const uint64_t* s4, uint64_t* out, int n) {
  orc_rng r;
  memcpy(r.s, s4, sizeof r.s);
  for (int i = 0; i < n; ++i) out[i] = xoshiro_next(&r);
}
ORC_EXPORT void orc_rng_draws(uint64_t seed, double* uni, int nu, double* nrm, int nn) {
  orc_rng r;
  rng_seed(&r, seed);
  for (int i = 0; i < nu; ++i) uni[i] = rng_uniform(&r);
  for (int i = 0; i < nn; ++i) nrm[i] = rng_normal(&r);
}

/*
 * Gaussian-vortex parameters: for k = 0..nvort-1, the centre is three
 * normals normalised to the sphere, then the amplitude is amp_scale times
 * one normal.  Output centres[nvort][3], amps[nvort].
 */
ORC_EXPORT void orc_gauss_params(uint64_t seed, int nvort, double amp_scale, double* centers, double* amps) {
  orc_rng r;
  rng_seed(&r, seed);
  for (int k = 0; k < nvort; ++k) {
    double c[3] = {rng_normal(&r), rng_normal(&r), rng_normal(&r)};
    g_normalize(c);
    centers[3 * k + 0] = c[0];
    centers[3 * k + 1] = c[1];
    centers[3 * k + 2] = c[2];
    amps[k] = amp_scale * rng_normal(&r);
  }
}

/*
 * Relative vorticity of the composite initial condition at n points:
 *   zeta = sbr * 2 * (2 pi) * z                         (SolidBodyRotation,
 *                                                         src/LpmVorticityGallery.hpp:25-40)
 *        + 2 w sin(lat) - K (n+1)(n+2) cos(lat)^n sin(lat) cos(n lon)   (RH wave, if rh_n > 0)
 *        + sum_k amps[k] exp(-(1 - x.c_k) / sigma^2)                    (Gaussian vortices)
 * with lat = atan2(z, sqrt(x^2+y^2)) and lon = atan4(y, x) as in
 * src/LpmGeometry.hpp:216-228.
 */
ORC_EXPORT void orc_ic_eval(const double* x, int n, double* zeta, double sbr, int rh_n, double rh_w, double rh_K,
                            int nvort, const double* centers, const double* amps, double sigma) {
#pragma omp parallel for
  for (int i = 0; i < n; ++i) {
    const double* p = &x[3 * i];
    double z = 0.0;
    if (sbr != 0.0) z += sbr * (2 * (2 * PI) * p[2]);
    if (rh_n > 0) {
      const double lat = g_latitude(p), lon = g_longitude(p);
      const double cl = cos(lat), sl = sin(lat);
      z += 2 * rh_w * sl - rh_K * (rh_n + 1) * (rh_n + 2) * pow(cl, rh_n) * sl * cos(rh_n * lon);
    }
    for (int k = 0; k < nvort; ++k) z += amps[k] * exp(-(1.0 - g_dot(p, &centers[3 * k])) / (sigma * sigma));
    zeta[i] = z;
  }
}

/* ------------------------------------------------------------------ */
/* error norms: src/LpmErrorNorms.cpp:26-70 (ErrNorms::compute, the       */
/* ENormScalar reducer of src/LpmErrorNorms.hpp:13-30)                 */
/* ------------------------------------------------------------------ */
/* Vector norms (:26-45): out = {l1, l2, linf} with e = appx - exact (:16-24),
 * l1 = sum |e| w / sum |x| w, l2 = sqrt(sum |e|^2 w / sum |x|^2 w),
 * linf = max |e| / max |x|. */
ORC_EXPORT void orc_err_norms_vec(const double* appx, const double* exact, const double* wt, int n, double* out) {
  double l1n = 0, l1d = 0, l2n = 0, l2d = 0, lin = -DBL_MAX, lid = -DBL_MAX;
  for (int i = 0; i < n; ++i) {
    double e[3] = {appx[3 * i] - exact[3 * i], appx[3 * i + 1] - exact[3 * i + 1], appx[3 * i + 2] - exact[3 * i + 2]};
    const double em = g_mag(e), xm = g_mag(&exact[3 * i]);
    l1n += em * wt[i];
    l1d += xm * wt[i];
    l2n += em * em * wt[i];
    l2d += xm * xm * wt[i];
    if (em > lin) lin = em;
    if (xm > lid) lid = xm;
  }
  out[0] = l1n / l1d;
  out[1] = sqrt(l2n / l2d);
  out[2] = lin / lid;
}
/* Scalar norms (:47-70): l1 and l2 as above with |e| and |x|; the maxima
 * are SIGNED (:64-65: linfnum = max e, linfdenom = max exact, starting
 * from the max-reduction identity -DBL_MAX), so an error that is negative
 * everywhere gives linf <= 0. */
ORC_EXPORT void orc_err_norms_scalar(const double* appx, const double* exact, const double* wt, int n, double* out) {
  double l1n = 0, l1d = 0, l2n = 0, l2d = 0, lin = -DBL_MAX, lid = -DBL_MAX;
  for (int i = 0; i < n; ++i) {
    const double e = appx[i] - exact[i], x = exact[i];
    l1n += fabs(e) * wt[i];
    l1d += fabs(x) * wt[i];
    l2n += (e * e) * wt[i];
    l2d += (x * x) * wt[i];
    lin = e > lin ? e : lin;
    lid = x > lid ? x : lid;
  }
  out[0] = l1n / l1d;
  out[1] = sqrt(l2n / l2d);
  out[2] = lin / lid;
}

/* Y_5^4 and its BVE velocity: src/LpmRossbyWaves.hpp:18-53 */
ORC_EXPORT void orc_sphharm54(const double* x, int n, double* f, double* u) {
  for (int i = 0; i < n; ++i) {
    const double* p = &x[3 * i];
    const double lam = g_longitude(p);
    const double z = p[2];
    f[i] = 30 * cos(4 * lam) * (z * (z * z - 1.0) * (z * z - 1.0));
    const double theta = g_latitude(p);
    const double coslat = cos(theta), sinlat = sin(theta), coslon = cos(lam), sinlon = sin(lam);
    const double usph = 0.5 * cos(4 * lam) * (coslat * coslat * coslat) * (5 * cos(2 * theta) - 3);
    const double vsph = 4 * (coslat * coslat * coslat) * sinlat * sin(4 * lam);
    u[3 * i + 0] = -usph * sinlon - vsph * sinlat * coslon;
    u[3 * i + 1] = usph * coslon - vsph * sinlat * sinlon;
    u[3 * i + 2] = vsph * coslat;
  }
}
