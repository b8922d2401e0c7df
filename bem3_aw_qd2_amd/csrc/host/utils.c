/*
 * utils.c -- 3-vector helpers, checked allocation and Gauss-Legendre tables.
 *
 * These are part of the ABI the reference exports to its callers
 * (com_src/my_utils.h, com_src/gauleg.h); example2.c allocates with m_alloc2.
 */
#include "bem3_aw_qd2.h"

void vadd_d(double *r, double *v0, double *v1)
{
  for (int i = 0; i < 3; i++) r[i] = v0[i] + v1[i];
}

void vsub_d(double *r, double *v0, double *v1)
{
  for (int i = 0; i < 3; i++) r[i] = v0[i] - v1[i];
}

double vabs_d(double *v)
{
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* normalise in place; -1 (and r unchanged) for the zero vector (my_utils.h:12-13) */
int vuni_d(double *r)
{
  double a = vabs_d(r);
  if (a == 0.0) return -1;
  r[0] /= a;
  r[1] /= a;
  r[2] /= a;
  return 0;
}

double vabs_2dm(double *v0, double *v1)
{
  double d0 = v0[0] - v1[0], d1 = v0[1] - v1[1], d2 = v0[2] - v1[2];
  return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

double vdot_d(double *v0, double *v1)
{
  return v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
}

void vcrs_d(double *r, double *v0, double *v1)
{
  r[0] = v0[1] * v1[2] - v0[2] * v1[1];
  r[1] = v0[2] * v1[0] - v0[0] * v1[2];
  r[2] = v0[0] * v1[1] - v0[1] * v1[0];
}

void prt_z_en(double complex z)
{
  printf("% 15.14e %+15.14eI\n", creal(z), cimag(z));
}

void prt_z_fen(char *txt, int n, double complex z)
{
  if (n >= 0) printf("%s[%3d] = % 15.14e %+15.14eI\n", txt, n, creal(z), cimag(z));
  else printf("%s = % 15.14e %+15.14eI\n", txt, creal(z), cimag(z));
}

/* zero-filled allocation that ends the process on failure, as the reference's
 * callers expect (com_src/my_utils.c:67-78) */
void *m_alloc2(size_t num, size_t size, char *txt)
{
  void *p = calloc(num, size);
  if (p == NULL) {
    printf("memory allocation error!\n%s. Exit..\n", txt ? txt : "");
    exit(1);
  }
  return p;
}

void continue_message(void)
{
  printf("continue? (y/n) : ");
  if (getchar() != 'y') {
    printf("Exit\n");
    exit(0);
  }
}

/*
 * Gauss-Legendre nodes and weights on [x1, x2] by Newton iteration on the
 * three-term Legendre recurrence, started from cos(pi (i - 1/4)/(n + 1/2)) and
 * stopped at |dz| <= 3e-16.  The reference builds its GLN/GHN tables with this
 * textbook procedure (com_src/gauleg.c:3-36); the same starting points,
 * recurrence and stopping rule are used so that the nodes agree bit for bit
 * (ex.dat stores the reference's tables, which the tests compare against), since
 * the element-rule switch compares a quadrature of dF/dv against a threshold.
 */
void gauleg(double x1, double x2, double x[], double w[], int n)
{
  const double tol = 3.0e-16;
  const int half = (n + 1) / 2;
  const double mid = 0.5 * (x2 + x1), len = 0.5 * (x2 - x1);
  for (int i = 1; i <= half; i++) {
    double z = cos(M_PI * ((double)i - 0.25) / ((double)n + 0.5));
    double pn, pn1, dp;
    for (;;) {
      double a = 1.0, b = 0.0; /* P_j, P_{j-1} */
      for (int j = 1; j <= n; j++) {
        double c = b;
        b = a;
        a = ((double)(2 * j - 1) * z * b - (double)(j - 1) * c) / (double)j;
      }
      pn = a;
      pn1 = b;
      dp = (double)n * (z * pn - pn1) / (z * z - 1.0);
      double zold = z;
      z = zold - pn / dp;
      if (fabs(z - zold) <= tol) break;
    }
    x[i - 1] = mid - len * z;
    x[n - i] = mid + len * z;
    w[i - 1] = 2.0 * len / ((1.0 - z * z) * dp * dp);
    w[n - i] = w[i - 1];
  }
  if (x1 == -x2 && n % 2 == 1) x[n / 2] = 0.0;
}
