// helixphm: a phantom program for phm2helix (the role of the reference's
// helical/dynphm.c): writes the phantom at acquisition view `iview` of a
// time-varying phantom whose elements drift sinusoidally with one period
// per rotation.
//
//   helixphm --base FILE [--period N] [--drift A] iview nview phmfile
//
// FILE holds one element per line: type cx cy u v rot_deg atten p0 p1 p2 p3
// (the phantom file format plus four drift phases).  At view v, with
// t = 2 pi v / N:  cx + A sin(t + p0), cy + A sin(t + p1),
// max(0.01, u + A sin(t + p2)), max(0.01, v + A sin(t + p3)); rotation and
// attenuation are fixed.  Numbers are written with 17 significant digits, so
// the phantom read back is the one computed.
#include <getopt.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char* const argv[]) {
  static const option opts[] = {{"base", 1, nullptr, 'b'},
                                {"period", 1, nullptr, 'p'},
                                {"drift", 1, nullptr, 'd'},
                                {nullptr, 0, nullptr, 0}};
  std::string base;
  double period = 1152, drift = 0.05;
  int c;
  while ((c = getopt_long(argc, argv, "", opts, nullptr)) != -1) {
    if (c == 'b') base = optarg;
    else if (c == 'p') period = atof(optarg);
    else if (c == 'd') drift = atof(optarg);
    else return 1;
  }
  if (optind + 3 != argc || base.empty()) {
    fprintf(stderr, "usage: %s --base FILE [--period N] [--drift A] iview nview phmfile\n", argv[0]);
    return 1;
  }
  const int view = atoi(argv[optind]);
  FILE* in = fopen(base.c_str(), "r");
  if (!in) {
    fprintf(stderr, "helixphm: cannot read %s\n", base.c_str());
    return 2;
  }
  FILE* out = fopen(argv[optind + 2], "w");
  if (!out) {
    fprintf(stderr, "helixphm: cannot write %s\n", argv[optind + 2]);
    fclose(in);
    return 2;
  }
  const double t = 2 * M_PI * view / period;
  char type[80];
  double cx, cy, u, v, rot, att, p[4];
  while (fscanf(in, "%79s %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", type, &cx, &cy, &u, &v, &rot,
                &att, &p[0], &p[1], &p[2], &p[3]) == 11) {
    const double ncx = cx + drift * sin(t + p[0]), ncy = cy + drift * sin(t + p[1]);
    const double nu = std::max(0.01, u + drift * sin(t + p[2]));
    const double nv = std::max(0.01, v + drift * sin(t + p[3]));
    fprintf(out, "%s %.17g %.17g %.17g %.17g %.17g %.17g\n", type, ncx, ncy, nu, nv, rot, att);
  }
  fclose(in);
  return fclose(out) == 0 ? 0 : 2;
}
