// ReadData.hxx -- file ingestion with CSRx normalisation, the interface of the
// reference's src/ReadData.hxx and the conversion routines of
// src/MATUtil.hxx (same names, arguments and return codes), implemented in
// faspxx_readdata.cpp over the drop-in MAT / VEC.
#pragma once

#include <vector>

#include "faspxx_dropin.hxx"

FaspRetCode ReadVEC(const char* fileName, VEC& dst);
FaspRetCode ReadMTX(const char* fileName, USI& row, USI& col, USI& nnz, std::vector<USI>& rowInd,
                    std::vector<USI>& colInd, std::vector<DBL>& values);
FaspRetCode ReadCSR(const char* fileName, USI& row, USI& col, USI& nnz, std::vector<USI>& rowPtr,
                    std::vector<USI>& colInd, std::vector<DBL>& values);
FaspRetCode ReadMat(const char* fileName, MAT& dst);
// The normalised CSRx arrays ReadMat hands to MAT (.csr or .mtx), on the host.
// No reference counterpart: ReadMat's steps without the device upload.
FaspRetCode ReadCSRx(const char* fileName, USI& row, USI& col, USI& nnz,
                     std::vector<USI>& rowPtr, std::vector<USI>& colInd,
                     std::vector<DBL>& values);

FaspRetCode SortCSRRow(const USI& row, const USI& col, const USI& nnz,
                       const std::vector<USI>& rowPtr, std::vector<USI>& colInd,
                       std::vector<DBL>& values);
FaspRetCode CSRtoMAT(const USI& row, const USI& col, const USI& nnz,
                     const std::vector<DBL>& values, const std::vector<USI>& colInd,
                     const std::vector<USI>& rowPtr, MAT& mat);
FaspRetCode MTXtoMAT(const USI& row, const USI& col, const USI& nnz,
                     const std::vector<USI>& rowInd, const std::vector<USI>& colInd,
                     const std::vector<DBL>& values, MAT& mat);
