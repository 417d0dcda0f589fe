/*
 * exception.h -- error types of the C++ skeleton layer.  The class names are
 * those of the reference (include/detail/exception.h:45-320) so that code
 * catching or constructing them compiles; here they are thrown (C++ API) or
 * reported as status codes (C ABI) instead of printed and ignored.
 */
#pragma once

#include <ostream>
#include <stdexcept>
#include <string>

namespace msl {
namespace detail {

class Exception : public std::runtime_error {
public:
    explicit Exception(const std::string &what) : std::runtime_error(what) {}
    Exception() : std::runtime_error("muesli exception") {}
    friend std::ostream &operator<<(std::ostream &os, const Exception &e) { return os << e.what(); }
};

// A failed C-ABI call (HIP error, invalid argument, allocation failure).
class RuntimeError : public Exception {
public:
    explicit RuntimeError(const std::string &what) : Exception(what) {}
};

class NotYetImplementedException : public Exception {
public:
    NotYetImplementedException() : Exception("not yet implemented") {}
};
class NotSameSizeException : public Exception {
public:
    NotSameSizeException() : Exception("not same size") {}
};
class IllegalDistributionException : public Exception {
public:
    IllegalDistributionException() : Exception("illegal distribution") {}
};
class UndefinedSourceException : public Exception {
public:
    UndefinedSourceException() : Exception("undefined source") {}
};
class UndefinedDestinationException : public Exception {
public:
    UndefinedDestinationException() : Exception("undefined destination") {}
};
class NonLocalAccessException : public Exception {
public:
    NonLocalAccessException() : Exception("non local access") {}
};
class MissingInitializationException : public Exception {
public:
    MissingInitializationException() : Exception("missing initialization") {}
};
class RotateRowCompleteNotImplementedException : public Exception {
public:
    RotateRowCompleteNotImplementedException() : Exception("rotate row complete not implemented") {}
};
class RotateColCompleteNotImplementedException : public Exception {
public:
    RotateColCompleteNotImplementedException() : Exception("rotate col complete not implemented") {}
};
class RotateRowManyNotImplementedException : public Exception {
public:
    RotateRowManyNotImplementedException() : Exception("rotate row many not implemented") {}
};
class IllegalGetException : public Exception {
public:
    IllegalGetException() : Exception("illegal get") {}
};
class IllegalPutException : public Exception {
public:
    IllegalPutException() : Exception("illegal put") {}
};
class IllegalPartitionException : public Exception {
public:
    IllegalPartitionException() : Exception("illegal partition") {}
};
class PartitioningImpossibleException : public Exception {
public:
    PartitioningImpossibleException() : Exception("partitioning impossible") {}
};
class IllegalPermuteException : public Exception {
public:
    IllegalPermuteException() : Exception("illegal permute") {}
};
class IllegalAllToAllException : public Exception {
public:
    IllegalAllToAllException() : Exception("illegal all to all") {}
};
class IllegalFunctorException : public Exception {
public:
    IllegalFunctorException() : Exception("illegal functor") {}
};
class FeatureNotSupportedByDeviceException : public Exception {
public:
    FeatureNotSupportedByDeviceException() : Exception("feature not supported by device") {}
};
class NoSolutionException : public Exception {
public:
    NoSolutionException() : Exception("no solution") {}
};
class InternalErrorException : public Exception {
public:
    InternalErrorException() : Exception("internal error") {}
};
class EmptyHeapException : public Exception {
public:
    EmptyHeapException() : Exception("empty heap") {}
};
class EmptyStackException : public Exception {
public:
    EmptyStackException() : Exception("empty stack") {}
};
class EmptyQueueException : public Exception {
public:
    EmptyQueueException() : Exception("empty queue") {}
};
class DeviceOutOfMemory : public Exception {
public:
    DeviceOutOfMemory() : Exception("device out of memory") {}
};
class IndexOutOfBoundsException : public Exception {
public:
    IndexOutOfBoundsException() : Exception("index out of bounds") {}
};
class SizeMismatchException : public Exception {
public:
    SizeMismatchException() : Exception("size mismatch") {}
};
class FoldToManyBlocks : public Exception {
public:
    FoldToManyBlocks() : Exception("fold to many blocks") {}
};
class UsingGPUFunctionOutsideGPU : public Exception {
public:
    UsingGPUFunctionOutsideGPU() : Exception("using g p u function outside g p u") {}
};
class InvalidCube : public Exception {
public:
    InvalidCube() : Exception("invalid cube") {}
};

}  // namespace detail
}  // namespace msl
