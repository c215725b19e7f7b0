// lsCOMP_xpcs -- command-line driver of the uint16 codec with the interface of
// the reference's lsCOMP/example/lsCOMP_xpcs.cpp:10-190 (see lscomp_cli.h).
#include "lscomp_cli.h"

int main(int argc, char *argv[]) { return cli_main<uint16_t>(argc, argv, "xpcs"); }
