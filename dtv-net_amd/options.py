"""Geometry and path options with the reference's names (options.py:1-25).

Reads ``options.yaml`` from the current working directory when present (the
reference's behaviour, options.py:3), otherwise the copy shipped next to
this file.  Loaded with ``yaml.safe_load``.
"""
import os

import yaml

_here = os.path.dirname(os.path.abspath(__file__))
_path = "./options.yaml" if os.path.exists("./options.yaml") else os.path.join(_here, "options.yaml")
with open(_path) as _f:
    y = yaml.safe_load(_f)

trainPath = os.path.join(y["dataPath"], y["subDir"]["train"])
validPath = os.path.join(y["dataPath"], y["subDir"]["valid"])
inputTrainData = os.path.join(y["dataPath"], y["subDir"]["trainInput"])
inputValidData = os.path.join(y["dataPath"], y["subDir"]["validInput"])
outputPath = os.path.join(y["dataPath"], "r.raw")
pretrain3 = y["pretrain3"]
checkpointPath = y["checkpointPath"]
debugPath = y["debug"]
standardAngleNum = y["standardAngleNum"]
standardDetectorSize = y["standardDetectorSize"]
standardVolumeSize = y["standardVolumeSize"]
standardSID = y["standardSID"]
standardSDD = y["standardSDD"]
beijingAngleNum = y["beijingAngleNum"]
beijingPlanes = y["beijingPlanes"]
beijingSubDetectorSize = y["beijingSubDetectorSize"]
beijingVolumeSize = y["beijingVolumeSize"]
beijingParameterRoot = y["beijingParameterRoot"]
beijingSID = y["beijingSID"]
beijingSDD = y["beijingSDD"]
sampleInterval = y["sampleInterval"]
