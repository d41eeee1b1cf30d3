// pathtracer: command-line front end (reference src/main.cu:11-53).
#include "../../include/wpt_host.h"

int main(int argc, char** argv) { return wpt_host_main(argc, argv); }
