/* sysgen.c -- random sequential insertion for the synthetic water boxes.
 *
 * Candidates come from the caller (numpy PCG64, so the sequence is seeded
 * and deterministic); this routine only accepts or rejects them in order
 * against the points already placed, under periodic minimum image in an
 * orthorhombic box, using a cell grid.  Returns the new number of points. */
#include <math.h>
#include <stdlib.h>

int rsa_place(const double *cand, int ncand, double *pts, int count, int target, const double *box, double dmin) {
  int nc[3];
  for (int d = 0; d < 3; d++) {
    nc[d] = (int)(box[d] / dmin);
    if (nc[d] < 1) nc[d] = 1;
  }
  long ncell = (long)nc[0] * nc[1] * nc[2];
  int *head = malloc(ncell * sizeof(int));
  int *next = malloc((size_t)(target > 0 ? target : 1) * sizeof(int));
  for (long c = 0; c < ncell; c++) head[c] = -1;
  double d2min = dmin * dmin;
  int ci[3];
#define CELL(p)                                                               \
  for (int d = 0; d < 3; d++) {                                               \
    double f = (p)[d] / box[d];                                               \
    f -= floor(f);                                                            \
    int q = (int)(f * nc[d]);                                                 \
    ci[d] = q >= nc[d] ? nc[d] - 1 : q;                                       \
  }
  for (int i = 0; i < count; i++) {
    CELL(pts + 3 * i);
    long c = ci[0] + (long)nc[0] * (ci[1] + (long)nc[1] * ci[2]);
    next[i] = head[c];
    head[c] = i;
  }
  for (int t = 0; t < ncand && count < target; t++) {
    const double *p = cand + 3 * t;
    CELL(p);
    int ok = 1;
    for (int dz = -1; dz <= 1 && ok; dz++)
      for (int dy = -1; dy <= 1 && ok; dy++)
        for (int dx = -1; dx <= 1 && ok; dx++) {
          int x = (ci[0] + dx + nc[0]) % nc[0], y = (ci[1] + dy + nc[1]) % nc[1], z = (ci[2] + dz + nc[2]) % nc[2];
          for (int j = head[x + (long)nc[0] * (y + (long)nc[1] * z)]; j >= 0; j = next[j]) {
            double r2 = 0.0;
            for (int d = 0; d < 3; d++) {
              double v = pts[3 * j + d] - p[d];
              v -= box[d] * nearbyint(v / box[d]);
              r2 += v * v;
            }
            if (r2 < d2min) { ok = 0; break; }
          }
        }
    if (!ok) continue;
    for (int d = 0; d < 3; d++) pts[3 * count + d] = p[d];
    long c = ci[0] + (long)nc[0] * (ci[1] + (long)nc[1] * ci[2]);
    next[count] = head[c];
    head[c] = count;
    count++;
  }
  free(head);
  free(next);
  return count;
}
