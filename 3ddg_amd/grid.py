"""Python mirror of the reference's Grid/Block call surface (Grid.h:29-76, Block.h:34-106) over the
C ABI. A Grid is one register of a device context; Block methods forward to dg3d_operate,
dg3d_axpby and dg3d_zero with the reference's argument meaning (`this` is written, `old` is read).
"""
from . import dg3d


class Block:
    """The single block of a Grid on this rank (Block.h:34)."""

    def __init__(self, ctx, reg):
        self.ctx = ctx
        self.reg = reg

    # Block.cpp:343-356
    def OperateUFromV(self, axis, mode, oldBlock):
        self.ctx.operate(dg3d.OP_UFROMV, axis, mode, oldBlock.reg, self.reg)

    # Block.cpp:389-402
    def OperateSponge(self, axis, mode, oldBlock):
        self.ctx.operate(dg3d.OP_SPONGE, axis, mode, oldBlock.reg, self.reg)

    # Block.cpp:358-371 (writes Q into `this`, reads rho' of oldBlock)
    def OperateQFromDRho(self, axis, mode, oldBlock):
        self.ctx.operate(dg3d.OP_QFROMDRHO, axis, mode, oldBlock.reg, self.reg)

    # Block.cpp:373-386
    def OperateDRhoFromQ(self, axis, mode, oldBlock):
        self.ctx.operate(dg3d.OP_DRHOFROMQ, axis, mode, oldBlock.reg, self.reg)

    # Block.cpp:327-340
    def selfWeightedAdd(self, other, selfScalar, otherScalar):
        self.ctx.axpby(self.reg, selfScalar, other.reg, otherScalar)

    # Block.cpp:297-310
    def setZero(self):
        self.ctx.zero(self.reg)


class Grid:
    """One state register (a Grid copy in driver.cpp:210-213)."""

    def __init__(self, ctx, reg, evalAdvect=True, evalDiff=False, evalSponge=True):
        self.ctx = ctx
        self.reg = reg
        self.mBlock = Block(ctx, reg)
        self.mEvalAdvect = evalAdvect
        self.mEvalDiff = evalDiff
        self.mEvalSponge = evalSponge
        self.mTime = 0.0

    def copy_from(self, other):
        """Grid copy constructor (Grid.cpp:78-174)."""
        self.ctx.copy(self.reg, other.reg)
        self.mTime = other.mTime

    def communicate(self):
        """Grid::communicate (Grid.cpp:487-565)."""
        self.ctx.halo(self.reg)

    def setZero(self):
        self.mBlock.setZero()

    def RK33(self, dt, oldGrid, k1, k2):
        """Grid::RK33 (Grid.cpp:642-902) through the fine-grained operators."""
        self.mTime += dt
        oldGrid.mTime += dt
        k1.setZero()
        k2.setZero()
        self.setZero()
        for src, dst in ((oldGrid, k1), (k1, k2), (k2, self)):
            src.communicate()
            if self.mEvalDiff:
                for axis in range(3):
                    src.mBlock.OperateQFromDRho(axis, 0, src.mBlock)
                src.communicate()
                for axis in range(3):
                    dst.mBlock.OperateDRhoFromQ(axis, 0, src.mBlock)
            if self.mEvalAdvect:
                for axis in range(3):
                    dst.mBlock.OperateUFromV(axis, 0, src.mBlock)
                    dst.mBlock.OperateUFromV(axis, 1, src.mBlock)
            if self.mEvalSponge:
                dst.mBlock.OperateSponge(0, 0, src.mBlock)
            if dst is k1:
                k1.mBlock.selfWeightedAdd(oldGrid.mBlock, dt, 1.0)
            elif dst is k2:
                k2.mBlock.selfWeightedAdd(k1.mBlock, dt * 0.25, 0.25)
                k2.mBlock.selfWeightedAdd(oldGrid.mBlock, 1.0, 0.75)
            else:
                self.mBlock.selfWeightedAdd(k2.mBlock, dt * (2.0 / 3.0), 2.0 / 3.0)
                self.mBlock.selfWeightedAdd(oldGrid.mBlock, 1.0, 1.0 / 3.0)
