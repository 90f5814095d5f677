// scout_mis: drop-in for the reference's bin/scout_mis (same 5 positional arguments,
// input formats and stdout lines); the repair runs in libscout_hip.
#include "driver.h"

int main(int argc, char** argv) { return scout::run_driver(argc, argv, scout::Problem::MIS); }
