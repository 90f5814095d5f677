// scout_coloring: drop-in for the reference's bin/scout_coloring (same 5 positional arguments,
// input formats and stdout lines); the repair runs in libscout_hip.
#include "driver.h"

int main(int argc, char** argv) { return scout::run_driver(argc, argv, scout::Problem::COLORING); }
