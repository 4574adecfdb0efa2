// Exports the reference's host oracle axhelmReference (axhelm/axhelmReference.cpp,
// compiled from the nekBench tree by oracle/Makefile) with C linkage, so the
// tests can load it with ctypes.  Test infrastructure only.
#include "params.h"
#include "axhelmReference.cpp"

extern "C" void ref_axhelmReference(int Nq, int numElements, double lambda, const double* ggeo,
                                    const double* D, const double* solIn, double* solOut)
{
  axhelmReference(Nq, numElements, lambda, ggeo, D, solIn, solOut);
}
