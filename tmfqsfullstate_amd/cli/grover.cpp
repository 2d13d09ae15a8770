// grover: Grover search with materialized diffusion (reference examples/grover.cpp).
#include "grover_main.h"

int main(int argc, char *argv[]) { return tmfqs_cli::groverMain(argc, argv, "grover", "normal", true); }
