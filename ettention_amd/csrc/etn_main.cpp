// The command-line reconstruction tool (frontend/src/main.cpp): every step
// lives in libetn.so behind etn_main.
#include "etn/etn.h"

int main(int argc, char* argv[]) { return etn_main(argc, argv); }
