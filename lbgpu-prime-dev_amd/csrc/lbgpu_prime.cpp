// lbgpu_prime: runs a reference input deck on the MI355X library (the
// reference's executable of the same name, src/lbgpu_prime.cu:92).
#include <stdio.h>

#include "lbm_io.h"

int main(int argc, char **argv) {
  setbuf(stdout, nullptr);  // unbuffered screen output, as there
  return lbm_deck_main(argc, argv);
}
