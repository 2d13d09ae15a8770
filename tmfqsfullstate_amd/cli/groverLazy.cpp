// groverLazy: Grover search with the lazy affine diffusion (reference examples/groverLazy.cpp).
#include "grover_main.h"

int main(int argc, char *argv[]) { return tmfqs_cli::groverMain(argc, argv, "groverLazy", "lazy", false); }
