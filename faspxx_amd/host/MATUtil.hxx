// MATUtil.hxx -- forwards to the drop-in's conversion routines (src/MATUtil.hxx).
#pragma once
#include "ReadData.hxx"
