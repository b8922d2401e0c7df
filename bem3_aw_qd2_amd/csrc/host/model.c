/*
 * model.c -- AQD2 model: input files, allocation, element constants, incident
 * boundary data, ex.dat read/write, point-in-domain classification.
 *
 * Host logic of the reference's d3qd2_src/bem3_aw_qd2.c, kept compatible where
 * files or callers see it: the same text formats (pw.txt, periodicity.txt,
 * medium.txt, gmsh 2.2 ASCII with each surface element listed twice), the same
 * 1-based pointer-to-pointer arrays, and the same ex.dat byte layout
 * (bem3_aw_qd2.c:185-381): a raw image of AQD2 followed by the arrays.
 */
#include <stddef.h>
#include "bem3_aw_qd2.h"
#include "host_internal.h"

/* ------------------------------------------------------------------ allocation */
static void alloc_nodes(BOUD *bd)
{
  bd->rn = (double **)m_alloc2(bd->Nn + 1, sizeof(double *), "model.c, alloc_nodes(), rn");
  for (int i = 0; i <= bd->Nn; i++) bd->rn[i] = (double *)m_alloc2(3, sizeof(double), "model.c, rn[i]");
}

static void free_nodes(BOUD *bd)
{
  for (int i = 0; i <= bd->Nn; i++) free(bd->rn[i]);
  free(bd->rn);
  bd->Nn = 0;
}

static void alloc_elements(BOUD *bd)
{
  int Ne = bd->Ne;
  bd->ed = (int **)m_alloc2(Ne + 1, sizeof(int *), "model.c, ed");
  bd->eni = (int **)m_alloc2(Ne + 1, sizeof(int *), "model.c, eni");
  bd->md = (int *)m_alloc2(Ne + 1, sizeof(int), "model.c, md");
  bd->sd = (int *)m_alloc2(Ne + 1, sizeof(int), "model.c, sd");
  bd->gd = (int *)m_alloc2(Ne + 1, sizeof(int), "model.c, gd");
  bd->cr = (double ***)m_alloc2(Ne + 1, sizeof(double **), "model.c, cr");
  bd->cw = (double ***)m_alloc2(Ne + 1, sizeof(double **), "model.c, cw");
  bd->ren = (double ***)m_alloc2(Ne + 1, sizeof(double **), "model.c, ren");
  bd->wen = (double ***)m_alloc2(Ne + 1, sizeof(double **), "model.c, wen");
  bd->Pi = (double complex **)m_alloc2(Ne + 1, sizeof(double complex *), "model.c, Pi");
  bd->dPi = (double complex **)m_alloc2(Ne + 1, sizeof(double complex *), "model.c, dPi");
  for (int i = 0; i <= Ne; i++) {
    bd->ed[i] = (int *)m_alloc2(4, sizeof(int), "model.c, ed[i]");
    bd->eni[i] = (int *)m_alloc2(4, sizeof(int), "model.c, eni[i]");
    bd->cr[i] = (double **)m_alloc2(3, sizeof(double *), "model.c, cr[i]");
    bd->cw[i] = (double **)m_alloc2(3, sizeof(double *), "model.c, cw[i]");
    for (int d = 0; d < 3; d++) {
      bd->cr[i][d] = (double *)m_alloc2(4, sizeof(double), "model.c, cr[i][d]");
      bd->cw[i][d] = (double *)m_alloc2(3, sizeof(double), "model.c, cw[i][d]");
    }
    bd->ren[i] = (double **)m_alloc2(4, sizeof(double *), "model.c, ren[i]");
    bd->wen[i] = (double **)m_alloc2(4, sizeof(double *), "model.c, wen[i]");
    for (int j = 0; j < 4; j++) {
      bd->ren[i][j] = (double *)m_alloc2(3, sizeof(double), "model.c, ren[i][j]");
      bd->wen[i][j] = (double *)m_alloc2(3, sizeof(double), "model.c, wen[i][j]");
    }
    bd->Pi[i] = (double complex *)m_alloc2(4, sizeof(double complex), "model.c, Pi[i]");
    bd->dPi[i] = (double complex *)m_alloc2(4, sizeof(double complex), "model.c, dPi[i]");
  }
}

static void free_elements(BOUD *bd)
{
  for (int i = 0; i <= bd->Ne; i++) {
    free(bd->ed[i]);
    free(bd->eni[i]);
    for (int d = 0; d < 3; d++) {
      free(bd->cr[i][d]);
      free(bd->cw[i][d]);
    }
    free(bd->cr[i]);
    free(bd->cw[i]);
    for (int j = 0; j < 4; j++) {
      free(bd->ren[i][j]);
      free(bd->wen[i][j]);
    }
    free(bd->ren[i]);
    free(bd->wen[i]);
    free(bd->Pi[i]);
    free(bd->dPi[i]);
  }
  free(bd->ed); free(bd->eni); free(bd->md); free(bd->sd); free(bd->gd);
  free(bd->cr); free(bd->cw); free(bd->ren); free(bd->wen); free(bd->Pi); free(bd->dPi);
  bd->Ne = 0;
}

/* element counts per domain and the per-domain solution arrays (bem3_aw_qd2.c:1025-1083) */
static void alloc_domains(AQD2 *qd)
{
  int *cnt = (int *)m_alloc2(qd->MN + 1, sizeof(int), "model.c, alloc_domains(), cnt");
  for (int i = 1; i <= qd->bd.Ne; i++) {
    cnt[qd->bd.md[i]]++;
    cnt[qd->bd.sd[i]]++;
  }
  qd->bd.sb = (SUBD *)m_alloc2(qd->MN + 1, sizeof(SUBD), "model.c, sb");
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &qd->bd.sb[d];
    int n = cnt[d];
    s->Ne = n;
    s->sid = (int *)m_alloc2(n + 1, sizeof(int), "model.c, sid");
    s->P = (double complex **)m_alloc2(n + 1, sizeof(double complex *), "model.c, P");
    s->dP = (double complex **)m_alloc2(n + 1, sizeof(double complex *), "model.c, dP");
    s->pv = (double complex ***)m_alloc2(n + 1, sizeof(double complex **), "model.c, pv");
    s->dpv = (double complex ***)m_alloc2(n + 1, sizeof(double complex **), "model.c, dpv");
    for (int j = 0; j <= n; j++) {
      s->P[j] = (double complex *)m_alloc2(4, sizeof(double complex), "model.c, P[j]");
      s->dP[j] = (double complex *)m_alloc2(4, sizeof(double complex), "model.c, dP[j]");
      s->pv[j] = (double complex **)m_alloc2(4, sizeof(double complex *), "model.c, pv[j]");
      s->dpv[j] = (double complex **)m_alloc2(4, sizeof(double complex *), "model.c, dpv[j]");
      for (int k = 0; k < 4; k++) {
        s->pv[j][k] = (double complex *)m_alloc2(3, sizeof(double complex), "model.c, pv[j][k]");
        s->dpv[j][k] = (double complex *)m_alloc2(3, sizeof(double complex), "model.c, dpv[j][k]");
      }
    }
  }
  free(cnt);
}

static void free_domains(AQD2 *qd)
{
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &qd->bd.sb[d];
    for (int j = 0; j <= s->Ne; j++) {
      free(s->P[j]);
      free(s->dP[j]);
      for (int k = 0; k < 4; k++) {
        free(s->pv[j][k]);
        free(s->dpv[j][k]);
      }
      free(s->pv[j]);
      free(s->dpv[j]);
    }
    free(s->sid); free(s->P); free(s->dP); free(s->pv); free(s->dpv);
  }
  free(qd->bd.sb);
}

/* signed element lists: +i where domain d is the element's main side, -i where
 * it is the sub side, in element order (bem3_aw_qd2.c:1085-1102) */
static void fill_domain_lists(AQD2 *qd)
{
  for (int d = 0; d <= qd->MN; d++) {
    int c = 1;
    for (int i = 1; i <= qd->bd.Ne; i++) {
      if (qd->bd.md[i] == d) qd->bd.sb[d].sid[c++] = i;
      else if (qd->bd.sd[i] == d) qd->bd.sb[d].sid[c++] = -i;
    }
  }
}

/* ------------------------------------------------------------------ input files */
static void bad_read(const char *fn, const char *what)
{
  printf("%s, failed to read the %s. exit...\n", fn, what);
  exit(1);
}

static void read_periodicity(char *fn, AQD2 *qd)
{
  char buf[256] = "";
  double a, b;
  int v;
  FILE *fp = fopen(fn, "rt");
  if (fp == NULL) {
    printf("Can not open the '%s' file. Exit...\n", fn);
    exit(1);
  }
  strncpy(qd->ped_fn, fn, sizeof(qd->ped_fn) - 1);
  if (fgets(buf, 256, fp) == NULL || fgets(buf, 256, fp) == NULL) bad_read("read_periodicity_data()", "line");
  if (fscanf(fp, "%d\n", &v) != 1) bad_read("read_periodicity_data()", "ps");
  qd->bd.ps = v;
  if (fscanf(fp, "%lf %lf", &a, &b) != 2) bad_read("read_periodicity_data()", "vd1");
  qd->bd.qd.vd1[0] = a;
  qd->bd.qd.vd1[1] = b;
  if (fscanf(fp, "%lf %lf", &a, &b) != 2) bad_read("read_periodicity_data()", "vd2");
  qd->bd.qd.vd2[0] = a;
  qd->bd.qd.vd2[1] = b;
  fclose(fp);
}

static void alloc_media(AQD2 *qd)
{
  int n = qd->MN + 1;
  qd->rho0 = (double *)m_alloc2(n, sizeof(double), "model.c, rho0");
  qd->c0 = (double *)m_alloc2(n, sizeof(double), "model.c, c0");
  qd->k0 = (double *)m_alloc2(n, sizeof(double), "model.c, k0");
  qd->K0 = (double *)m_alloc2(n, sizeof(double), "model.c, K0");
  qd->k1 = (double complex *)m_alloc2(n, sizeof(double complex), "model.c, k1");
  qd->k2 = (double complex *)m_alloc2(n, sizeof(double complex), "model.c, k2");
}

static void read_media(char *fn, AQD2 *qd)
{
  char buf[256] = "";
  double a;
  int v;
  FILE *fp = fopen(fn, "rt");
  if (fp == NULL) {
    printf("Can not open the '%s' file. Exit...\n", fn);
    exit(1);
  }
  strncpy(qd->med_fn, fn, sizeof(qd->med_fn) - 1);
  if (fgets(buf, 256, fp) == NULL || fgets(buf, 256, fp) == NULL) bad_read("read_medium_data()", "line");
  if (fscanf(fp, "%d\n", &v) != 1) bad_read("read_medium_data()", "MN");
  qd->MN = v;
  alloc_media(qd);
  if (fgets(buf, 256, fp) == NULL) bad_read("read_medium_data()", "line");
  for (int i = 1; i <= qd->MN; i++) {
    if (fscanf(fp, "%lf", &a) != 1) bad_read("read_medium_data()", "rho0");
    qd->rho0[i] = a;
    if (fscanf(fp, "%lf", &a) != 1) bad_read("read_medium_data()", "c0");
    qd->c0[i] = a;
  }
  fclose(fp);
}

static int scan_int(FILE *fp, const char *what)
{
  int v;
  if (fscanf(fp, "%d", &v) != 1) bad_read("read_mesh_data()", what);
  return v;
}

static int domain_tag(int tag, AQD2 *qd)
{
  if (tag == OPENDID) tag = 0;
  if (tag > qd->MN) {
    printf("domain id %d is not defined medium data. check domain and medium data. exit..\n", tag);
    exit(1);
  }
  return tag;
}

/*
 * gmsh 2.2 ASCII mesh (bem3_aw_qd2.c:531-828).  Every surface element appears
 * twice: id 2i-1 carries the main-side domain (first tag; 99 is the open
 * lattice domain 0) and the node order, id 2i the opposite side with its node
 * order reversed (checked).  When the main side is the open domain the two
 * sides are exchanged so that domain 0 is always a sub side's partner... more
 * precisely: elements facing the open region get md = 0 with reversed nodes.
 */
static void read_mesh(char *fn, AQD2 *qd)
{
  char buf[256] = "";
  double a;
  FILE *fp = fopen(fn, "rt");
  if (fp == NULL) {
    printf("Can not open the '%s' file. Exit...\n", fn);
    exit(1);
  }
  strncpy(qd->msh_fn, fn, sizeof(qd->msh_fn) - 1);
  if (fgets(buf, 256, fp) == NULL) bad_read("read_mesh_data()", "line");
  if (fscanf(fp, "%lf", &a) != 1) bad_read("read_mesh_data()", "msh version");
  if (a < 2.2) {
    printf("This program supports mesh file version %g later. Reading file version is %g. Exit...\n", 2.2, a);
    fclose(fp);
    exit(1);
  }
  if (scan_int(fp, "msh format") != 0) {
    printf("This program supports 'ASCII' data format mesh file. Exit...\n");
    fclose(fp);
    exit(1);
  }
  int prec;
  if (fscanf(fp, "%d\n", &prec) != 1) bad_read("read_mesh_data()", "data precision");
  if (prec != 8) {
    printf("This program supports double precision mesh data. Exit...\n");
    fclose(fp);
    exit(1);
  }
  if (fgets(buf, 256, fp) == NULL || fgets(buf, 256, fp) == NULL) bad_read("read_mesh_data()", "line");
  int nn;
  if (fscanf(fp, "%d\n", &nn) != 1) bad_read("read_mesh_data()", "Nn");
  qd->bd.Nn = nn;
  alloc_nodes(&qd->bd);
  for (int i = 1; i <= nn; i++) {
    int id = scan_int(fp, "id");
    if (id != i) printf("bad id %d\n", id);
    for (int d = 0; d < 3; d++) {
      if (fscanf(fp, d < 2 ? "%lf" : "%lf\n", &a) != 1) bad_read("read_mesh_data()", "rn");
      qd->bd.rn[i][d] = a;
    }
  }
  if (fgets(buf, 256, fp) == NULL || fgets(buf, 256, fp) == NULL) bad_read("read_mesh_data()", "line");
  qd->bd.Ne = scan_int(fp, "Ne") / 2;
  alloc_elements(&qd->bd);

  int nc = 0;
  BOUD *bd = &qd->bd;
  for (int i = 1; i <= bd->Ne; i++) {
    /* first listing: main side */
    if (scan_int(fp, "id") != 2 * i - 1) {
      printf("bad id. Exit...\n");
      exit(1);
    }
    int et = scan_int(fp, "type");
    if (et != ELT3 && et != ELT4) {
      printf("bad element type. element type must be %d or %d. Exit...\n", ELT3, ELT4);
      exit(1);
    }
    int ntag = scan_int(fp, "number of tags");
    for (int j = 0; j < ntag; j++) {
      int tag = scan_int(fp, "physical entity");
      if (j == 0) bd->md[i] = domain_tag(tag, qd);
      else if (j == 1) bd->gd[i] = tag;
    }
    bd->ed[i][0] = scan_int(fp, "ed[i][0]");
    bd->ed[i][1] = scan_int(fp, "ed[i][1]");
    bd->ed[i][2] = scan_int(fp, "ed[i][2]");
    bd->ed[i][3] = (et == ELT3) ? 0 : scan_int(fp, "ed[i][3]");
    int ncorner = (et == ELT3) ? 3 : 4;
    for (int j = 0; j < 4; j++) bd->eni[i][j] = (j < ncorner) ? nc++ : -1;

    /* second listing: sub side, node order reversed */
    if (scan_int(fp, "element id") != 2 * i) {
      printf("bad id. Exit...\n");
      exit(1);
    }
    et = scan_int(fp, "element type");
    if (et != ELT3 && et != ELT4) {
      printf("bad element type. element type must be %d or %d. Exit...\n", ELT3, ELT4);
      exit(1);
    }
    ntag = scan_int(fp, "number of tags");
    for (int j = 0; j < ntag; j++) {
      int tag = scan_int(fp, "domain id");
      if (j == 0) bd->sd[i] = domain_tag(tag, qd);
    }
    static const int rev3[3] = {0, 2, 1}, rev4[4] = {0, 3, 2, 1};
    for (int j = 0; j < ((et == ELT3) ? 3 : 4); j++) {
      int v = scan_int(fp, "ed");
      int expect = (et == ELT3) ? bd->ed[i][rev3[j]] : bd->ed[i][rev4[j]];
      if (v != expect) {
        printf("node id miss matched. check element id %d. Exit...\n", i * 2);
        exit(1);
      }
    }
    /* the open domain always takes the main side */
    if (bd->sd[i] == 0) {
      bd->sd[i] = bd->md[i];
      bd->md[i] = 0;
      if (et == ELT3) {
        int t = bd->ed[i][1];
        bd->ed[i][1] = bd->ed[i][2];
        bd->ed[i][2] = t;
      } else {
        int t = bd->ed[i][3];
        bd->ed[i][3] = bd->ed[i][1];
        bd->ed[i][1] = t;
      }
    }
  }
  fclose(fp);
  bd->NN = nc;
}

void read_aqd2(int argc, char **argv, AQD2 *qd)
{
  if (argc != 6 && argc != 13) {
    printf("This program needs command line arguments as follows.\n");
    printf("%s plane_wave_datafile_name periodicity_datafile_name medium_datafile_name mesh_datafile_name output_datafile_name\n", argv[0]);
    printf("[rv_x rv_y rv_z theta tr_x tr_y tr_z](optional)\n");
    printf("rv : vector defining rotation axis, theta : rotation angle ( using Rodrigues' rotation formula ), tr : translation vector\n");
    exit(0);
  }
  read_data_apw(argv[1], &qd->pw);
  read_periodicity(argv[2], qd);
  read_media(argv[3], qd);
  read_mesh(argv[4], qd);
  if (argc == 13) {
    for (int i = 0; i < 3; i++) qd->rv[i] = atof(argv[6 + i]);
    qd->th = atof(argv[9]);
    for (int i = 0; i < 3; i++) qd->tv[i] = atof(argv[10 + i]);
  } else {
    qd->rv[0] = 1.0; qd->rv[1] = 0.0; qd->rv[2] = 0.0;
    qd->th = 0.0;
    qd->tv[0] = qd->tv[1] = qd->tv[2] = 0.0;
  }
}

void print_aqd2(AQD2 *qd)
{
  print_data_apw(&qd->pw);
  printf("--- periodicity data ---\n");
  printf("periodicity data file name = %s\n", qd->ped_fn);
  printf("periodicity setting        = %10d\n", qd->bd.ps);
  printf("lattice constant d1        = (% 7g,% 7g)\n", qd->bd.qd.vd1[0], qd->bd.qd.vd1[1]);
  printf("lattice constant d2        = (% 7g,% 7g)\n", qd->bd.qd.vd2[0], qd->bd.qd.vd2[1]);
  printf("\n");
  printf("--- medium data ---\n");
  printf("medium data file name      = %s\n", qd->med_fn);
  for (int i = 1; i <= qd->MN; i++) {
    printf("domain id %2d rho0 [kg/m^3] =%8.6g\n", i, qd->rho0[i]);
    printf("             c0      [m/s] =%8.6g\n", qd->c0[i]);
  }
  printf("\n");
  printf("--- mesh data ---\n");
  printf("mesh data file name        = %s\n", qd->msh_fn);
  printf("node number                = %8d\n", qd->bd.Nn);
  printf("defined element number     = %8d\n", qd->bd.Ne * 2);
  printf("\n");
  if (qd->th != 0.0 || vabs_d(qd->tv) != 0.0) {
    printf("-- rotation and translation settings --\n");
    if (qd->th != 0.0) {
      printf("vector defining rotation axis :(% 8.7g,% 8.7g,% 8.7g)\n", qd->rv[0], qd->rv[1], qd->rv[2]);
      printf("rotation angle           [rad]: %8.7g\n", qd->th);
    }
    if (vabs_d(qd->tv) != 0.0) printf("translation vector            :(%8.7g,%8.7g,%8.7g)\n", qd->tv[0], qd->tv[1], qd->tv[2]);
  }
  printf("\n");
}

/* ------------------------------------------------------------------ set-up */
/* Rodrigues rotation about rv by th, then translation by tv (bem3_aw_qd2.c:394-422) */
static void move_object(AQD2 *qd)
{
  double n[3] = { qd->rv[0], qd->rv[1], qd->rv[2] };
  vuni_d(n);
  double s = sin(qd->th), c = cos(qd->th), M[9];
  M[0] = c + n[0] * n[0] * (1.0 - c); M[1] = n[0] * n[1] * (1.0 - c) - n[2] * s; M[2] = n[2] * n[0] * (1.0 - c) + n[1] * s;
  M[3] = n[0] * n[1] * (1.0 - c) + n[2] * s; M[4] = c + n[1] * n[1] * (1.0 - c); M[5] = n[1] * n[2] * (1.0 - c) - n[0] * s;
  M[6] = n[2] * n[0] * (1.0 - c) - n[1] * s; M[7] = n[1] * n[2] * (1.0 - c) + n[0] * s; M[8] = c + n[2] * n[2] * (1.0 - c);
  for (int i = 1; i <= qd->bd.Nn; i++) {
    double *r = qd->bd.rn[i], t[3];
    for (int d = 0; d < 3; d++) t[d] = M[3 * d] * r[0] + M[3 * d + 1] * r[1] + M[3 * d + 2] * r[2] + qd->tv[d];
    for (int d = 0; d < 3; d++) r[d] = t[d];
  }
}

/*
 * Element geometry constants (bem3_aw_qd2.c:941-987): bilinear quads use
 * r(z,e) = c0 + c1 z + c2 e + c3 z e from the four corners and the normal
 * w = dr/dz x dr/de = cw0 + cw1 z + cw2 e; triangles use the equilateral
 * reference triangle with corners (-1/2,-sqrt3/2), (1,0), (-1/2,sqrt3/2) so that
 * r = c0 + c1 z + c2 e and the constant normal cw0.
 */
void aqd2_element_constants(BOUD *bd)
{
  for (int i = 1; i <= bd->Ne; i++) {
    double rc[3][4];
    for (int d = 0; d < 3; d++)
      for (int j = 0; j < 4; j++) rc[d][j] = (bd->ed[i][j] != 0 || j < 3) ? bd->rn[bd->ed[i][j]][d] : bd->rn[0][d];
    double **cr = bd->cr[i], **cw = bd->cw[i];
    if (bd->ed[i][3] != 0) {
      for (int d = 0; d < 3; d++) {
        cr[d][0] = 0.25 * (rc[d][0] + rc[d][1] + rc[d][2] + rc[d][3]);
        cr[d][1] = 0.25 * (-rc[d][0] + rc[d][1] + rc[d][2] - rc[d][3]);
        cr[d][2] = 0.25 * (-rc[d][0] - rc[d][1] + rc[d][2] + rc[d][3]);
        cr[d][3] = 0.25 * (rc[d][0] - rc[d][1] + rc[d][2] - rc[d][3]);
      }
      /* component d of w uses the two other coordinates (a, b) cyclically */
      for (int d = 0; d < 3; d++) {
        int a = (d + 1) % 3, b = (d + 2) % 3;
        cw[d][0] = 0.125 * ((rc[a][0] - rc[a][2]) * (rc[b][1] - rc[b][3]) - (rc[a][1] - rc[a][3]) * (rc[b][0] - rc[b][2]));
        cw[d][1] = 0.125 * ((rc[a][2] - rc[a][3]) * (rc[b][0] - rc[b][1]) - (rc[a][0] - rc[a][1]) * (rc[b][2] - rc[b][3]));
        cw[d][2] = 0.125 * ((rc[a][1] - rc[a][2]) * (rc[b][0] - rc[b][3]) - (rc[a][0] - rc[a][3]) * (rc[b][1] - rc[b][2]));
      }
    } else {
      for (int d = 0; d < 3; d++) {
        cr[d][0] = 1.0 / 3.0 * (rc[d][0] + rc[d][1] + rc[d][2]);
        cr[d][1] = 1.0 / 3.0 * (-rc[d][0] + 2.0 * rc[d][1] - rc[d][2]);
        cr[d][2] = 1.0 / sqrt(3.0) * (rc[d][2] - rc[d][0]);
        cr[d][3] = 0.0;
      }
      for (int d = 0; d < 3; d++) {
        int a = (d + 1) % 3, b = (d + 2) % 3;
        cw[d][0] = (rc[a][1] - rc[a][0]) * (rc[b][2] - rc[b][0]) - (rc[b][1] - rc[b][0]) * (rc[a][2] - rc[a][0]);
        cw[d][1] = 0.0;
        cw[d][2] = 0.0;
      }
    }
  }
  aqd2_quadrature_tables(bd);
}

/* collocation-node and quadrature tables (bem3_aw_qd2.c:989-1023) */
void aqd2_quadrature_tables(BOUD *bd)
{
  const double a44 = 0.5773502691896257645091;           /* 1/sqrt3 */
  const double a49 = 0.7745966692414833770359;           /* sqrt(3/5) */
  const double w0 = 0.3086419753086419753086, w1 = 0.4938271604938271604938, w2 = 0.7901234567901234567901;
  static const double s44z[4] = {-1, 1, 1, -1}, s44e[4] = {-1, -1, 1, 1};
  for (int j = 0; j < 4; j++) {
    bd->zt_44[j] = s44z[j] * a44;
    bd->et_44[j] = s44e[j] * a44;
    bd->wt_44[j] = 1.0;
  }
  static const double z49[9] = {-1, 1, 1, -1, 0, 1, 0, -1, 0}, e49[9] = {-1, -1, 1, 1, -1, 0, 1, 0, 0};
  for (int j = 0; j < 9; j++) {
    bd->zt_49[j] = z49[j] * a49;
    bd->et_49[j] = e49[j] * a49;
    bd->wt_49[j] = (j < 4) ? w0 : ((j < 8) ? w1 : w2);
  }
  const double n0 = 0.2, n1 = 0.3464101615137754587055;
  bd->zt_34[0] = -n0; bd->zt_34[1] = 2.0 * n0; bd->zt_34[2] = -n0; bd->zt_34[3] = 0.0;
  bd->et_34[0] = -n1; bd->et_34[1] = 0.0; bd->et_34[2] = n1; bd->et_34[3] = 0.0;
  bd->wt_34[0] = bd->wt_34[1] = bd->wt_34[2] = 0.5208333333333333333333;
  bd->wt_34[3] = -0.5625;
  const double m0 = 0.3480702390148154917985, m1 = 0.2052130961576726346557;
  const double m2 = 0.6028753385763033130539, m3 = 0.3554395089236065568357;
  bd->zt_37[0] = -m0; bd->zt_37[1] = 2.0 * m0; bd->zt_37[2] = -m0;
  bd->zt_37[3] = m1; bd->zt_37[4] = -2.0 * m1; bd->zt_37[5] = m1; bd->zt_37[6] = 0.0;
  bd->et_37[0] = -m2; bd->et_37[1] = 0.0; bd->et_37[2] = m2;
  bd->et_37[3] = m3; bd->et_37[4] = 0.0; bd->et_37[5] = -m3; bd->et_37[6] = 0.0;
  bd->wt_37[0] = bd->wt_37[1] = bd->wt_37[2] = 0.1259391805448271525957;
  bd->wt_37[3] = bd->wt_37[4] = bd->wt_37[5] = 0.1323941527885061807376;
  bd->wt_37[6] = 0.225;
  gauleg(-1.0, 1.0, bd->xli, bd->wli, GLN);
  gauleg(-1.0, 1.0, bd->xhi, bd->whi, GHN);
}

/* collocation nodes, their normals and the incident potential there:
 * Pi = -p/k2, dPi = -(dp/dn)/k2 (bem3_aw_qd2.c:1104-1146) */
static void incident_node_data(AQD2 *qd)
{
  BOUD *bd = &qd->bd;
  double complex cp = -1.0 / qd->pw.wd.k2;
  for (int i = 1; i <= bd->Ne; i++) {
    int tri = (bd->ed[i][3] == 0);
    for (int j = 0; j < 4; j++) {
      double z = tri ? bd->zt_34[j] : bd->zt_44[j];
      double e = tri ? bd->et_34[j] : bd->et_44[j];
      double r[3], w[3];
      aqd2_rw(bd, i, z, e, r, w);
      for (int d = 0; d < 3; d++) {
        bd->ren[i][j][d] = r[d];
        bd->wen[i][j][d] = w[d];
      }
      vuni_d(w);
      double complex p, dp;
      calc_apw_dpdn(&p, &dp, r, w, &qd->pw);
      bd->Pi[i][j] = cp * p;
      bd->dPi[i][j] = cp * dp;
    }
  }
}

void initialize_aqd2(AQD2 *qd)
{
  setup_apw(&qd->pw);
  qd->bd.qd.k = qd->pw.wd.k0;
  qd->bd.qd.vk[0] = qd->pw.wd.kx;
  qd->bd.qd.vk[1] = qd->pw.wd.ky;
  qd->bd.qd.vk[2] = qd->pw.wd.kz;
  if (qd->bd.ps != 0) {
    /* the Bloch phase needs both in-plane components (bem3_aw_qd2.c:77-84) */
    if (qd->bd.qd.vk[0] == 0.0 || qd->bd.qd.vk[1] == 0.0) {
      printf("kx and ky must not be equal to 0 (must satisfy quasi-peridic boundary condition). Exit...\n");
      exit(1);
    }
    if (qd->bd.qd.vk[2] == 0.0)
      printf("kz must not be equal to 0 (The wave source and the objcet must not overlap). Exit...\n");
  }
  qd->rho0[0] = qd->pw.rho0;
  qd->c0[0] = qd->pw.c0;
  qd->k0[0] = qd->pw.wd.k0;
  qd->K0[0] = qd->pw.wd.K0;
  qd->k1[0] = qd->pw.wd.k1;
  qd->k2[0] = qd->pw.wd.k2;
  for (int i = 1; i <= qd->MN; i++) {
    qd->k0[i] = qd->pw.wd.omega / qd->c0[i];
    qd->K0[i] = qd->rho0[i] * qd->c0[i] * qd->c0[i];
    qd->k1[i] = I * qd->pw.wd.omega / qd->K0[i];
    qd->k2[i] = I * qd->pw.wd.omega * qd->rho0[i];
  }
  if (qd->th != 0.0 || vabs_d(qd->tv) != 0.0) move_object(qd);
  aqd2_element_constants(&qd->bd);
  alloc_domains(qd);
  fill_domain_lists(qd);
  incident_node_data(qd);
}

void finalize_aqd2(AQD2 *qd)
{
  aqd2_release_device(qd);
  free_domains(qd);
  free_elements(&qd->bd);
  free_nodes(&qd->bd);
  free(qd->rho0); free(qd->c0); free(qd->k0); free(qd->K0); free(qd->k1); free(qd->k2);
}

/* ------------------------------------------------------------------ ex.dat */
static void wr(const void *p, size_t sz, size_t n, FILE *fp)
{
  if (fwrite(p, sz, n, fp) != n) {
    printf("dat_write_aqd2(), write error. exit...\n");
    exit(1);
  }
}

static void rd(void *p, size_t sz, size_t n, FILE *fp, const char *what)
{
  if (fread(p, sz, n, fp) != n) {
    printf("bem3_aw_qd2.c, dat_read_qqd2(), failed to read the %s. exit...\n", what);
    exit(1);
  }
}

/* byte layout of bem3_aw_qd2.c:185-224 */
void dat_write_aqd2(char *fname, AQD2 *qd)
{
  FILE *fp = fopen(fname, "wb");
  if (fp == NULL) {
    printf("dat_write_aqd2(), Failed to create the %s file.\n", fname);
    exit(1);
  }
  BOUD *bd = &qd->bd;
  int n = qd->MN + 1;
  wr(qd, sizeof(AQD2), 1, fp);
  wr(qd->rho0, sizeof(double), n, fp);
  wr(qd->c0, sizeof(double), n, fp);
  wr(qd->k0, sizeof(double), n, fp);
  wr(qd->K0, sizeof(double), n, fp);
  wr(qd->k1, sizeof(double complex), n, fp);
  wr(qd->k2, sizeof(double complex), n, fp);
  wr(&qd->pw, sizeof(Apw), 1, fp);
  for (int i = 0; i <= bd->Nn; i++) wr(bd->rn[i], sizeof(double), 3, fp);
  for (int i = 0; i <= bd->Ne; i++) wr(bd->ed[i], sizeof(int), 4, fp);
  for (int i = 0; i <= bd->Ne; i++) wr(bd->eni[i], sizeof(int), 4, fp);
  wr(bd->md, sizeof(int), bd->Ne + 1, fp);
  wr(bd->sd, sizeof(int), bd->Ne + 1, fp);
  wr(bd->gd, sizeof(int), bd->Ne + 1, fp);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) wr(bd->ren[i][j], sizeof(double), 3, fp);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) wr(bd->wen[i][j], sizeof(double), 3, fp);
  for (int i = 0; i <= bd->Ne; i++) wr(bd->Pi[i], sizeof(double complex), 4, fp);
  for (int i = 0; i <= bd->Ne; i++) wr(bd->dPi[i], sizeof(double complex), 4, fp);
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &bd->sb[d];
    wr(&s->Ne, sizeof(int), 1, fp);
    wr(s->sid, sizeof(int), s->Ne + 1, fp);
    for (int i = 0; i <= s->Ne; i++) wr(s->P[i], sizeof(double complex), 4, fp);
    for (int i = 0; i <= s->Ne; i++) wr(s->dP[i], sizeof(double complex), 4, fp);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(s->pv[i][j], sizeof(double complex), 3, fp);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(s->dpv[i][j], sizeof(double complex), 3, fp);
  }
  fclose(fp);
}

void dat_read_aqd2(char *fname, AQD2 *qd)
{
  FILE *fp = fopen(fname, "rb");
  if (fp == NULL) {
    printf("dat_read_aqd2(), Failed to open the %s file.\n", fname);
    exit(1);
  }
  rd(qd, sizeof(AQD2), 1, fp, "qd");
  /* the image holds the writer's pointers; none of them is used */
  int n = qd->MN + 1;
  alloc_media(qd);
  rd(qd->rho0, sizeof(double), n, fp, "rho0");
  rd(qd->c0, sizeof(double), n, fp, "c0");
  rd(qd->k0, sizeof(double), n, fp, "k0");
  rd(qd->K0, sizeof(double), n, fp, "K0");
  rd(qd->k1, sizeof(double complex), n, fp, "k1");
  rd(qd->k2, sizeof(double complex), n, fp, "k2");
  rd(&qd->pw, sizeof(Apw), 1, fp, "pw");
  BOUD *bd = &qd->bd;
  alloc_nodes(bd);
  alloc_elements(bd);
  for (int i = 0; i <= bd->Nn; i++) rd(bd->rn[i], sizeof(double), 3, fp, "rn[i]");
  for (int i = 0; i <= bd->Ne; i++) rd(bd->ed[i], sizeof(int), 4, fp, "ed[i]");
  for (int i = 0; i <= bd->Ne; i++) rd(bd->eni[i], sizeof(int), 4, fp, "eni[i]");
  rd(bd->md, sizeof(int), bd->Ne + 1, fp, "md");
  rd(bd->sd, sizeof(int), bd->Ne + 1, fp, "sd");
  rd(bd->gd, sizeof(int), bd->Ne + 1, fp, "gd");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(bd->ren[i][j], sizeof(double), 3, fp, "ren[i][j]");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(bd->wen[i][j], sizeof(double), 3, fp, "wen[i][j]");
  for (int i = 0; i <= bd->Ne; i++) rd(bd->Pi[i], sizeof(double complex), 4, fp, "Pi[i]");
  for (int i = 0; i <= bd->Ne; i++) rd(bd->dPi[i], sizeof(double complex), 4, fp, "dPi[i]");
  aqd2_element_constants(bd);
  alloc_domains(qd);
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &bd->sb[d];
    int tmp;
    rd(&tmp, sizeof(int), 1, fp, "tmp");
    rd(s->sid, sizeof(int), s->Ne + 1, fp, "sid");
    for (int i = 0; i <= s->Ne; i++) rd(s->P[i], sizeof(double complex), 4, fp, "P[i]");
    for (int i = 0; i <= s->Ne; i++) rd(s->dP[i], sizeof(double complex), 4, fp, "dP[i]");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(s->pv[i][j], sizeof(double complex), 3, fp, "pv[i][j]");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(s->dpv[i][j], sizeof(double complex), 3, fp, "dpv[i][j]");
  }
  fclose(fp);
}

/* ------------------------------------------------------------------ geometry helpers */
void aqd2_rw(const BOUD *bd, int s, double z, double e, double *r, double *w)
{
  double **cr = bd->cr[s], **cw = bd->cw[s];
  if (bd->ed[s][3] != 0) {
    for (int d = 0; d < 3; d++) {
      r[d] = cr[d][0] + cr[d][1] * z + cr[d][2] * e + cr[d][3] * z * e;
      w[d] = cw[d][0] + cw[d][1] * z + cw[d][2] * e;
    }
  } else {
    for (int d = 0; d < 3; d++) {
      r[d] = cr[d][0] + cr[d][1] * z + cr[d][2] * e;
      w[d] = cw[d][0];
    }
  }
}

void domain_id_l_aqd2(int *l1, int *l2, double *rt, AQD2 *qd)
{
  const double *d1 = qd->bd.qd.vd1, *d2 = qd->bd.qd.vd2;
  double idd = 1.0 / (d1[0] * d2[1] - d2[0] * d1[1]);
  *l1 = (int)floor((rt[0] * d2[1] - rt[1] * d2[0]) * idd + 0.5);
  *l2 = (int)floor(-(rt[0] * d1[1] - rt[1] * d1[0]) * idd + 0.5);
}

/* sizes and offsets of the public structs as compiled here, for binding checks:
 * sizeof(AQD2), offsetof(AQD2, bd), offsetof(BOUD, sb), sizeof(Apw),
 * sizeof(QPD2), offsetof(AQD2, MN) */
int aqd2_layout(long *out, int n)
{
  long v[6] = {(long)sizeof(AQD2), (long)offsetof(AQD2, bd), (long)offsetof(BOUD, sb), (long)sizeof(Apw),
               (long)sizeof(QPD2), (long)offsetof(AQD2, MN)};
  int m = n < 6 ? n : 6;
  for (int i = 0; i < m; i++) out[i] = v[i];
  return m;
}
