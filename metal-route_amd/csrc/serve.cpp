// Serving loop: filled in by the continuous-batching implementation.
#include "serve.h"
