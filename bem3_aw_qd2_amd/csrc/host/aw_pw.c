/*
 * aw_pw.c -- incident plane wave: input file, set-up and field values.
 *
 * Same file format, rotation convention and field expressions as the
 * reference's d3qd2_src/aw_pw.c:3-199 (pw.txt is shared with it).  These are
 * O(1) per point; the GPU kernels carry their own device copy of the same
 * expressions (device_pw in the HIP sources) for points evaluated there.
 */
#include "bem3_aw_qd2.h"

static void die_line(const char *what)
{
  printf("aw_pw.c, read_data_apw(), failed to read the %s. exit...\n", what);
  exit(1);
}

/* pw.txt: two comment lines, "rho0 c0 f", one comment line,
 * "Re(p) Im(p) tx ty tz theta psi" (aw_pw.c:3-77) */
void read_data_apw(char *fname, Apw *pw)
{
  FILE *fp = fopen(fname, "rt");
  char buf[256] = "";
  double a, b;
  if (fp == NULL) {
    printf("failed to read %s. no beams defined. Exit...\n", fname);
    exit(1);
  }
  if (fgets(buf, 256, fp) == NULL || fgets(buf, 256, fp) == NULL) die_line("line");
  if (fscanf(fp, "%lf", &a) != 1) die_line("rho0");
  pw->rho0 = a;
  if (fscanf(fp, "%lf", &a) != 1) die_line("c0");
  pw->c0 = a;
  if (fscanf(fp, "%lf\n", &a) != 1) die_line("f");
  pw->f = a;
  if (fgets(buf, 256, fp) == NULL) die_line("line");
  if (fscanf(fp, "%lf", &a) != 1) die_line("real(p)");
  if (fscanf(fp, "%lf", &b) != 1) die_line("imag(p)");
  pw->p = a + I * b;
  for (int i = 0; i < 3; i++) {
    if (fscanf(fp, "%lf", &a) != 1) die_line("tv");
    pw->tv[i] = a;
  }
  if (fscanf(fp, "%lf", &a) != 1) die_line("theta");
  pw->theta = a;
  if (fscanf(fp, "%lf", &a) != 1) die_line("psi");
  pw->psi = a;
  fclose(fp);
}

void print_data_apw(Apw *pw)
{
  double absI = 0.5 * creal(pw->p * conj(pw->p)) / (pw->rho0 * pw->c0);
  printf("-- planewave --\n");
  printf("medium density                   [kg/m^3] : %15.14g\n", pw->rho0);
  printf("speed of sound in the medium        [m/s] : %15.14g\n", pw->c0);
  printf("frequency                            [Hz] : %15.14g\n", pw->f);
  printf("sound pressure amplitude             [Pa] :%7.6g+%7.6gI\n", creal(pw->p), cimag(pw->p));
  printf("x-component of translation vector     [m] : %15.14g\n", pw->tv[0]);
  printf("y-component of translation vector     [m] : %15.14g\n", pw->tv[1]);
  printf("z-component of translation vector     [m] : %15.14g\n", pw->tv[2]);
  printf("rotation parameter theta            [rad] : %15.14g\n", pw->theta);
  printf("rotation parameter psi              [rad] : %15.14g\n", pw->psi);
  printf("-- additional information --\n");
  printf("wavelength                            [m] : %15.14g\n", pw->c0 / pw->f);
  printf("absolute value of sound intensity [W/m^2] : %15.14g\n", absI);
  printf("x-component of sound intensity    [W/m^2] : %15.14g\n", absI * sin(pw->theta) * cos(pw->psi));
  printf("y-component of sound intensity    [W/m^2] : %15.14g\n", absI * sin(pw->theta) * sin(pw->psi));
  printf("z-component of sound intensity    [W/m^2] : %15.14g\n", absI * cos(pw->theta));
  printf("\n");
}

/* derived constants and the rotation taking the z axis onto the propagation
 * direction (theta from z, psi the azimuth), aw_pw.c:103-126 */
void setup_apw(Apw *pw)
{
  double ct = cos(pw->theta), st = sin(pw->theta), cp = cos(pw->psi), sp = sin(pw->psi);
  double *R = pw->wd.R;
  pw->wd.omega = 2.0 * M_PI * pw->f;
  pw->wd.K0 = pw->c0 * pw->c0 * pw->rho0;
  pw->wd.k1 = I * pw->wd.omega / pw->wd.K0;
  pw->wd.k2 = I * pw->wd.omega * pw->rho0;
  pw->wd.k0 = pw->wd.omega / pw->c0;
  pw->wd.lambda0 = pw->c0 / pw->f;
  R[0] = ct * cp * cp + sp * sp; R[1] = sp * cp * (ct - 1.0); R[2] = st * cp;
  R[3] = sp * cp * (ct - 1.0);   R[4] = ct * sp * sp + cp * cp; R[5] = st * sp;
  R[6] = -st * cp;               R[7] = -st * sp;               R[8] = ct;
  pw->wd.kx = R[2] * pw->wd.k0;
  pw->wd.ky = R[5] * pw->wd.k0;
  pw->wd.kz = R[8] * pw->wd.k0;
}

/* coordinates in the beam frame (aw_pw.c:138-141) */
static void to_beam_frame(double *r0, const double *r, const Apw *pw)
{
  const double *R = pw->wd.R;
  double rc[3] = { r[0] - pw->tv[0], r[1] - pw->tv[1], r[2] - pw->tv[2] };
  r0[0] = R[0] * rc[0] + R[1] * rc[1] - R[2] * rc[2];
  r0[1] = R[3] * rc[0] + R[4] * rc[1] - R[5] * rc[2];
  r0[2] = -R[6] * rc[0] - R[7] * rc[1] + R[8] * rc[2];
}

void calc_apw_pv(double complex *p, double complex *v, double *r, Apw *pw)
{
  double r0[3];
  to_beam_frame(r0, r, pw);
  double arg = pw->wd.k0 * r0[2];
  double complex pp = pw->p * (cos(arg) + I * sin(arg));
  double complex vz = pp / (pw->rho0 * pw->c0);
  const double *R = pw->wd.R;
  *p = pp;
  /* beam-frame velocity is (0, 0, vz); rotate back */
  for (int i = 0; i < 3; i++) v[i] = R[3 * i + 0] * 0.0 + R[3 * i + 1] * 0.0 + R[3 * i + 2] * vz;
}

void calc_apw_dpdn(double complex *p, double complex *dpdn, double *r, double *n, Apw *pw)
{
  const double *R = pw->wd.R;
  double r0[3];
  to_beam_frame(r0, r, pw);
  double n0z = -R[6] * n[0] - R[7] * n[1] + R[8] * n[2];
  double arg = pw->wd.k0 * r0[2];
  double complex pp = pw->p * (cos(arg) + I * sin(arg));
  *p = pp;
  *dpdn = pp * I * pw->wd.k0 * n0z;
}
