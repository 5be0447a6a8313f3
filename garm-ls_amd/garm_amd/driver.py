"""Time-step rule of the reference driver.

Simulator::advanceTimeBySteps (Simulator.cpp:102-128) with
LevelSetLiquid::getTimeStep = cfl * dx / |u|max (LevelSetLiquid.h:129): the
step is clamped to the end of the frame and the last two steps of a frame
share the remaining time.  |u|max = 0 gives an infinite step, which the
clamp turns into the rest of the frame.
"""
import math


class FrameClock:
    def __init__(self, dx, cfl, frame_rate=25):
        self.dx = dx
        self.cfl = cfl
        self.target = 1.0 / frame_rate
        self.time = 0.0

    def next_dt(self, max_speed):
        dt = self.cfl * self.dx / max_speed if max_speed > 0 else math.inf
        if self.time + dt >= self.target:
            dt = self.target - self.time
            self.time = 0.0          # frame done; the next call starts a new frame
            return dt
        if self.time + 2 * dt >= self.target:
            dt = (self.target - self.time) / 2
        self.time += dt
        return dt
