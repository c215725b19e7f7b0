// lsCOMP_cssi -- command-line driver of the uint32 codec with the interface of
// the reference's lsCOMP/example/lsCOMP_cssi.cpp:10-190 (see lscomp_cli.h).
#include "lscomp_cli.h"

int main(int argc, char *argv[]) { return cli_main<uint32_t>(argc, argv, "cssi"); }
