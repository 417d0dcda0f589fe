/*
 * argtype.h -- additional functor arguments (reference include/argtype.h):
 * objects registered with FunctorBase::addArgument are told the tile width
 * and asked to refresh before a skeleton runs.
 */
#pragma once

namespace msl {

class ArgumentType {
public:
    virtual ~ArgumentType() {}
    virtual void update() {}
    virtual int getSmemSize() const { return 0; }
    virtual void setTileWidth(int value) { tile_width = value; }

protected:
    int tile_width = -1;
};

}  // namespace msl
