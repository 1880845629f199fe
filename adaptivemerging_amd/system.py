"""Python mirror of the reference's Java call surface over libam.

`RigidBodySystem`, `CollisionProcessor` and `PGS` keep the names, fields and
call order of src/mergingBodies3D/{RigidBodySystem,CollisionProcessor,PGS}.java
so code (and tests) written against the reference read the same.  Every
method runs on the GPU through the C ABI; nothing here computes physics.
"""
from __future__ import annotations

import time

import numpy as np

from ._lib import AM_SPRING_ZERO
from .context import Context


class PGS:
    """m3d/PGS.java:15-63: fields of the projected Gauss-Seidel solver."""

    def __init__(self, iterations: int = 200, ctx: Context | None = None):
        self.ctx = ctx
        self.changes: list[float] = []
        self.init(iterations)

    def init(self, iterations: int):
        """PGS.init (PGS.java:25-34)."""
        self.iterations = iterations
        self.tolerance = 1e-5
        self.omega = 1.0
        self.feedbackStiffness = 0.0
        self.compliance = 0.0
        self.contacts = None
        self.warmStart = False
        self.computeInCollection = False
        self.postStabilization = False

    def _apply(self, ctx: Context, restitutionOverride, restitutionOverrideVal, frictionOverride, frictionOverrideVal):
        p = ctx.params
        if self.computeInCollection:
            p.iterations_in_collection = self.iterations
        else:
            p.iterations = self.iterations
        p.tolerance = self.tolerance
        p.omega = self.omega
        p.feedback_stiffness = self.feedbackStiffness
        p.enable_post_stabilization = 0
        p.enable_compliance = 1 if self.compliance != 0.0 else 0
        p.compliance = self.compliance
        p.restitution_override = int(restitutionOverride)
        p.restitution = restitutionOverrideVal
        p.friction_override = int(frictionOverride)
        p.friction = frictionOverrideVal
        ctx.set_params()

    def solve(self, dt: float, restitutionOverride=False, restitutionOverrideVal=0.5, frictionOverride=False,
              frictionOverrideVal=0.1):
        """PGS.solve (PGS.java:73-194) over the context's current contacts."""
        ctx = self.ctx
        self._apply(ctx, restitutionOverride, restitutionOverrideVal, frictionOverride, frictionOverrideVal)
        if self.computeInCollection:
            return ctx.update_in_collections(dt)
        it, _ = ctx.solve(dt, self.postStabilization)
        self.changes = list(ctx.changes())
        return it


class CollisionProcessor:
    """m3d/CollisionProcessor.java: detection, warm start, solve drivers and
    the parameters of :1058-1083 (same names, same defaults)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.solver = PGS(ctx=ctx)
        self.collisionDetectTime = 0.0
        self.collisionSolveTime = 0.0
        self.collectionUpdateTime = 0.0
        self.contactOrderingTime = 0.0
        self.singleItPGSTime = 0.0
        self.badWarmStarts = 0
        self.badWarmStartsRepaired = 0
        # parameters (CollisionProcessor.java:1058-1083); the Java field
        # `warmStart` shares its name with the method, so it is useWarmStart here
        self.useWarmStart = True
        self.shuffle = False
        self.feedbackStiffness = 0.5
        self.enablePostStabilization = False
        self.enableCompliance = True
        self.compliance = 1e-3
        self.restitutionOverride = False
        self.restitution = 0.5
        self.frictionOverride = False
        self.friction = 0.1
        self.iterations = 30
        self.tolerance = 1e-5
        self.omega = 1.0
        self.iterationsInCollection = 1
        self._n = 0

    @property
    def contacts(self) -> list[dict]:
        """The current contact list (CollisionProcessor.contacts, :45)."""
        c = self.ctx.contacts()
        keys = list(c.keys())
        return [{k: c[k][i] for k in keys} for i in range(len(c["body1"]))]

    def collisionDetection(self, dt: float):
        """collisionDetection (:91-102)."""
        t0 = time.perf_counter()
        self._n = self.ctx.detect(dt)
        self.collisionDetectTime = time.perf_counter() - t0

    def updateContactsMap(self):
        """updateContactsMap (:443-448)."""
        self.ctx.update_contacts_map()

    def updateBodyPairContacts(self):
        """updateBodyPairContacts (:145-164): body-pair groups are rebuilt on
        the device with every detection; nothing to do on the host."""

    def warmStart(self, postStabilization: bool = False):
        """warmStart (:481-640)."""
        self.badWarmStarts, self.badWarmStartsRepaired = self.ctx.warm_start(postStabilization)

    def redoWarmStart(self):
        """redoWarmStart (:453-459)."""
        self.ctx.redo_warm_start()

    def noWarmStart(self):
        """noWarmStart (:464-470)."""
        self.ctx.no_warm_start()

    def solveLCP(self, dt: float, postStabilization: bool = False):
        """solveLCP (:108-137): configures the solver as the reference does."""
        if self._n == 0:
            self.collisionSolveTime = 0.0
            return 0
        self.solver.init(self.iterations)
        self.solver.tolerance = self.tolerance
        self.solver.omega = self.omega
        self.solver.feedbackStiffness = (self.feedbackStiffness if (not self.enablePostStabilization or
                                                                    postStabilization) else 0.0)
        self.solver.postStabilization = postStabilization
        self.solver.compliance = self.compliance if self.enableCompliance else 0.0
        self.solver.warmStart = True
        t0 = time.perf_counter()
        it = self.solver.solve(dt, self.restitutionOverride, self.restitution, self.frictionOverride,
                               self.friction)
        self.collisionSolveTime = time.perf_counter() - t0
        return it

    def clearBodyPairContacts(self):
        """clearBodyPairContacts (:213-226) and the body-pair histories."""
        self.ctx.clear_body_pair_contacts()

    def updateInCollections(self, dt: float, mergeParams=None):
        """updateInCollections (:232-303): one sweep inside collections."""
        t0 = time.perf_counter()
        p = self.ctx.params
        p.iterations_in_collection = self.iterationsInCollection
        p.feedback_stiffness = self.feedbackStiffness
        p.enable_compliance = int(self.enableCompliance)
        p.compliance = self.compliance
        self.ctx.set_params()
        self.ctx.update_in_collections(dt)
        self.collectionUpdateTime = time.perf_counter() - t0


class Merging:
    """m3d/Merging.java: merge / unmerge over the context's body pairs, with
    the MergeParameters names and defaults (:31-48)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.mergingEvent = False
        self.mergingBuildTime = 0.0     # MergeParameters.mergingBuildTime
        self.unmergeAllRequested = False  # MergeParameters.unmergeAll (the "unmerge all" request)
        self.unmergingBuildTime = 0.0
        self.enableMerging = True
        self.enableMergePinned = True
        self.enableMergeStableContactCondition = True
        self.enableMergeLetItBreathe = True
        self.enableUnmerging = True
        self.enableUnmergeFrictionCondition = True
        self.enableUnmergeNormalCondition = True
        self.enableUnmergeRelativeMotionCondition = True
        self.updateContactsInCollections = True
        self.stepAccumMerging = 3
        self.stepAccumUnmerging = 3
        self.stepsBetweenMergeEvents = 10
        self.thresholdMerge = 1e-2
        self.thresholdUnmerge = 2e-2
        self.thresholdBreath = 1e-5
        self.metricPositionLevel = False  # MergeParameters.metricPositionLevel

    def apply(self):
        p = self.ctx.params
        p.enable_merging = int(self.enableMerging)
        p.enable_merge_pinned = int(self.enableMergePinned)
        p.enable_merge_stable_contact = int(self.enableMergeStableContactCondition)
        p.enable_merge_let_it_breathe = int(self.enableMergeLetItBreathe)
        p.enable_unmerging = int(self.enableUnmerging)
        p.enable_unmerge_friction = int(self.enableUnmergeFrictionCondition)
        p.enable_unmerge_normal = int(self.enableUnmergeNormalCondition)
        p.enable_unmerge_relative_motion = int(self.enableUnmergeRelativeMotionCondition)
        p.update_contacts_in_collections = int(self.updateContactsInCollections)
        p.step_accum_merging = self.stepAccumMerging
        p.step_accum_unmerging = self.stepAccumUnmerging
        p.steps_between_merge_events = self.stepsBetweenMergeEvents
        p.threshold_merge = self.thresholdMerge
        p.threshold_unmerge = self.thresholdUnmerge
        p.threshold_breath = self.thresholdBreath
        p.metric_position_level = int(self.metricPositionLevel)
        self.ctx.set_params()

    def unmergeAll(self) -> int:
        """Merging.unmergeAll (:168-206); clears the request."""
        self.unmergeAllRequested = False
        return self.ctx.unmerge_all()

    def merge(self) -> int:
        """Merging.merge (:73-163)."""
        n = self.ctx.merge()
        self.mergingEvent = self.mergingEvent or n > 0
        return n

    def unmerge(self, dt: float) -> int:
        """Merging.unmerge (:215-273)."""
        n = self.ctx.unmerge(dt)
        self.mergingEvent = self.mergingEvent or n > 0
        return n


class Sleeping:
    """m3d/Sleeping.java: sleep / wake over the context's bodies, with the
    SleepParameters names and defaults (:25-30)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.enableSleeping = True
        self.stepAccum = 10
        self.threshold = 1e-5
        self.wakeAllRequested = False   # SleepParameters.wakeAll (the "wake all" button)

    def apply(self):
        p = self.ctx.params
        p.enable_sleeping = int(self.enableSleeping)
        p.sleep_step_accum = self.stepAccum
        p.sleep_threshold = self.threshold
        self.ctx.set_params()

    def wake(self) -> int:
        """Sleeping.wake (:107-139)."""
        return self.ctx.wake()

    def sleep(self) -> int:
        """Sleeping.sleep (:49-98); returns the top-level bodies asleep."""
        return self.ctx.sleep()

    def wakeAll(self):
        """Sleeping.wakeAll (:144-150)."""
        self.ctx.wake_all()
        self.wakeAllRequested = False


class RigidBodySystem:
    """m3d/RigidBodySystem.java: the per-step phase order of advanceTime
    (:102-185)."""

    def __init__(self, scene, device: int = 0):
        self.scene = scene
        self.ctx = Context(device)
        self.ctx.upload(scene)
        self.collision = CollisionProcessor(self.ctx)
        self.merging = Merging(self.ctx)
        self.sleeping = Sleeping(self.ctx)
        self.useGravity = True
        self.gravityAmount = 1.0
        self.gravityAngle = 90.0
        self.springStiffnessMod = 1.0   # RigidBodySystem.java:553
        self.springDampingMod = 1.0     # RigidBodySystem.java:554
        self.globalViscousLinearDecay = 1.0   # RigidBodySystem.java:550
        self.globalViscousAngularDecay = 1.0  # RigidBodySystem.java:551
        self.totalSteps = 0
        self.simulationTime = 0.0
        self.computeTime = 0.0
        self.warmStartTime = 0.0
        self.mergingTime = 0.0
        self.unmergingTime = 0.0
        # CSV export (RigidBodySystem.java:495-549): off by default
        self.saveCSV = False
        self.sceneName = "scene"
        self._csv = None

    @property
    def bodies(self):
        return self.scene.bodies

    def applyExternalForces(self, dt: float):
        """applyExternalForces (:234-257): gravity (:276-290), then the
        scene's springs (applySpringForces :309-322)."""
        p = self.ctx.params
        p.use_gravity = int(self.useGravity)
        p.gravity_amount = self.gravityAmount
        p.gravity_angle_deg = self.gravityAngle
        p.spring_stiffness_mod = self.springStiffnessMod
        p.spring_damping_mod = self.springDampingMod
        p.viscous_linear_decay = self.globalViscousLinearDecay
        p.viscous_angular_decay = self.globalViscousAngularDecay
        self.ctx.set_params()
        self.ctx.apply_gravity()
        self.ctx.apply_spring_forces()
        self.ctx.apply_impulses()

    def applyImpulse(self, body: int, pointW, force):
        """Holds a force at a world point for the next step's external-force
        pass (the impulse tool, RigidBodySystem.applyImpulse :262-267)."""
        self.ctx.impulse_add(body, pointW, force)

    def advanceTime(self, dt: float):
        """advanceTime (:102-185)."""
        t0 = time.perf_counter()
        self.totalSteps += 1
        self.merging.apply()
        self.sleeping.apply()
        self.ctx.clear_bodies()
        # prescribed motions where the animate tool acts (:109)
        self.ctx.kinematic_update(self.simulationTime)
        if self.sleeping.wakeAllRequested:
            self.sleeping.wakeAll()
        if self.merging.unmergeAllRequested:
            self.merging.unmergeAll()
        self.applyExternalForces(dt)
        self.collision.updateContactsMap()
        self.collision.collisionDetection(dt)
        self.collision.updateBodyPairContacts()
        t1 = time.perf_counter()
        self.collision.warmStart(False)
        self.warmStartTime = time.perf_counter() - t1
        self.sleeping.wake()
        if self.merging.updateContactsInCollections:
            self.collision.updateInCollections(dt)
        t1 = time.perf_counter()
        self.ctx.accumulate_for_unmerging(dt)
        if self.merging.unmerge(dt):
            self.merging.unmergingBuildTime = time.perf_counter() - t1
        else:
            self.merging.unmergingBuildTime = 0.0
        if self.merging.mergingEvent:
            self.ctx.clear_top_level()
        self.unmergingTime = time.perf_counter() - t1
        if self.collision.useWarmStart:
            self.collision.redoWarmStart()
        else:
            self.collision.noWarmStart()
        self.collision.solveLCP(dt, False)
        self.collision.clearBodyPairContacts()
        self.ctx.advance_velocities(dt)
        t1 = time.perf_counter()
        self.ctx.accumulate_for_merging(dt)
        self.mergingTime = time.perf_counter() - t1
        self.ctx.advance_positions(dt)
        if self.collision.enablePostStabilization:
            self.postStabilization(dt)
        self.merging.mergingBuildTime = 0.0
        if self.totalSteps % self.merging.stepsBetweenMergeEvents == 0:
            t1 = time.perf_counter()
            if self.merging.merge():
                self.merging.mergingBuildTime = time.perf_counter() - t1
            self.mergingTime += time.perf_counter() - t1
        self.sleeping.sleep()
        self.ctx.apply_viscous_decay()
        self.computeTime = time.perf_counter() - t0
        self.simulationTime += dt
        self.ctx.sim_time = self.simulationTime
        self.exportDataToFile()

    def postStabilization(self, dt: float):
        """postStabilization (:354-378): clear, detect again, warm start and
        solve with the post flag, move the positions by deltaV.  Runs with
        merging and unmerging off, as am_step does."""
        if self.merging.enableMerging or self.merging.enableUnmerging:
            raise ValueError("post-stabilization runs with merging and unmerging off")
        self.ctx.clear_bodies()
        self.collision.updateContactsMap()
        self.collision.collisionDetection(dt)
        self.collision.warmStart(True)
        self.collision.solveLCP(dt, True)
        self.ctx.advance_positions_post(dt)

    def exportDataToFile(self):
        """exportDataToFile (:495-542): while saveCSV is on, the first call
        opens <sceneName>.csv (<sceneName>_merged.csv with merging on) and
        writes the header, later calls one line per step; turning saveCSV off
        closes the file.  contactOrdering is 0: contacts are not reordered
        breadth first here; singleItPGS is the whole single sweep."""
        from .csv_export import CSVExporter

        if self.saveCSV:
            if self._csv is None:
                name = self.sceneName + ("_merged.csv" if self.merging.enableMerging else ".csv")
                self._csv = CSVExporter(name)
            else:
                c = self.collision
                self._csv.row([len(self.ctx.structure()[1]), self.ctx.n_contacts(), c.collisionDetectTime,
                               self.warmStartTime, c.collisionSolveTime, c.collectionUpdateTime,
                               c.contactOrderingTime, c.collectionUpdateTime, self.mergingTime,
                               self.merging.mergingBuildTime, self.unmergingTime, self.merging.unmergingBuildTime,
                               self.computeTime])
        elif self._csv is not None:
            self._csv.close()
            self._csv = None

    def reset(self):
        """reset (:390-426): collections dissolved, bodies back to their
        initial poses and velocities, the contact history cleared, and each
        spring reset (Spring.reset, Spring.java:106-130): a zero spring's
        anchor moves to its point's initial world position."""
        self.ctx.upload(self.scene)
        springs = self.scene.spring_array().copy()
        for sp in springs:
            if int(sp["type"]) == AM_SPRING_ZERO:
                sp["pw"] = self.scene.body_point_world(int(sp["body1"]), sp["pb1"])
        if len(springs):
            self.ctx.set_springs(springs)
        self.totalSteps = 0
        self.simulationTime = 0.0
        self.ctx.sim_time = 0.0

    def state(self):
        x, R, v, w = self.ctx.pull_state()
        return {"x": x, "R": R.reshape(-1, 3, 3), "v": v, "omega": w}

    def kinetic_energy(self) -> float:
        b, _, _, _, _, _ = self.scene.arrays()
        x, R, v, w = self.ctx.pull_state()
        return float(0.5 * np.sum(b["mass"] * np.sum(v * v, axis=1)))
