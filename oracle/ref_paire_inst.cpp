// Compiles the reference's scOOP/mc/paire.cpp from where it lies and adds an
// explicit instantiation of EPatch<WcaTruncSq>, so that the harness can call
// EPatch::minDistSegments (defined in that .cpp) on raw inputs.
#include "mc/paire.cpp"
template class EPatch<WcaTruncSq>;
