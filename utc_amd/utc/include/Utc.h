// Utc.h -- umbrella header of the UTC task surface (CPU part).
#pragma once

#include "Barrier.h"
#include "ProcList.h"
#include "Task.h"
#include "TaskBase.h"
#include "Timer.h"
#include "UserTaskBase.h"
#include "UtcBasics.h"
#include "UtcContext.h"
