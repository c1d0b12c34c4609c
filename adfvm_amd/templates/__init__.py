"""Case files for the apps (the reference's templates/ directory): each
defines primal, objective, perturb, parameters, nSteps, writeInterval,
startTime and dt, and writes its generated case directory on first use."""
