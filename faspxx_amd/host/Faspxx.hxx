// Faspxx.hxx -- the reference header name, forwarding to the drop-in classes.
#pragma once
#include "faspxx_dropin.hxx"
