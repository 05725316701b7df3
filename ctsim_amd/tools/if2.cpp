// if2 --comp: Herman's comparative statistics d, r, e of two .if images,
// the first being the reference (tools/if2.cpp --comp, imagefile.cpp:185-240).
#include <cstdio>
#include <cstring>
#include <iostream>

#include "ctsim/ctsim.h"

int main(int argc, char* argv[]) {
  bool comp = false;
  const char* files[2] = {nullptr, nullptr};
  int nf = 0;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--comp")) comp = true;
    else if (argv[i][0] != '-' && nf < 2) files[nf++] = argv[i];
    else {
      std::cerr << "usage: " << argv[0] << " infile1 infile2 --comp\n";
      return 1;
    }
  }
  if (nf != 2 || !comp) {
    std::cerr << "usage: " << argv[0] << " infile1 infile2 --comp\n";
    return 1;
  }
  ImageFile a, b;
  if (!a.fileRead(files[0]) || !b.fileRead(files[1])) {
    std::cerr << "Unable to read image files\n";
    return 1;
  }
  double d, r, e;
  if (!a.comparativeStatistics(b, d, r, e)) return 1;
  std::cout << "d=" << d << ", r=" << r << ", e=" << e << std::endl;  // if2.cpp:220
  return 0;
}
